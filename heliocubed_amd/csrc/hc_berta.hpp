// Berta WENOZ fourth-order operator (spatial_operator 2) of the HelioCubed
// MHD RHS on MI355X: the kernels of one RK stage over a box table. Included
// by hc_o4.hip inside its anonymous namespace; shares its box table, scratch
// layout, metric tables, face-product helpers and Roe kernel.
//
// Restates the host path of BoxOp_EulerCubedSphere::BertaOperator
// (src/BoxOp_BertaCubedSphereImplem.H:1737-2076; src/ = HelioCubed reference
// repository) with the Jameson detector and Roe 8-wave dissipation:
//   sphAvgToPrim (:278-340)          WAvg = V(<U>) on box+5   (k_o4_wbar)
//   BertaComputeThetaJameson (:346-507)  raw flag on box+4, dilated on box+3
//   BertaAverageToPoint (:696-741) + sphPointToPrim (:212-275)  point W on box+4
//   per direction:
//     BertaReconstructWithFallback (:747-1239): WENOZ / WENO3 / PLM face points
//     Roe repair with the cell-average donors (:1817-1844)
//     point fluxes of both sides summed, theta-aware face averages (:1246-1266, :1904-1946)
//     BertaPrimPointToMappedFlux (:1334-1534) with the separable-metric shortcut (:1-98),
//       evaluated inside the flux kernel
//     primToMappedRoeFlux with theta-aware dissipation average (src/BoxOp_RoeCubedSphere.H:29-145),
//     or BertaPrimPointToMappedRusanovFlux (:1541-1730) for riemann_solver 1
//     F = (Sum - D)/2 (:2004-2026)
// Max/min follow the reference's std::max / std::min ((a < b) comparisons).

// Scratch slots of the Berta operator (the standard operator's, reused).
constexpr int F_WA = F_WB;                  // V(<U>), box+5
constexpr int F_THR = F_MEB;                // raw Jameson flag, box+4
constexpr int F_THETA = F_ETA;              // dilated flag, box+5 (0 outside box+3)
constexpr int F_WPT = F_W;                  // point primitives, box+4
constexpr int F_BWL = F_PTL, F_BWR = F_PTR; // face points (k_o4_roe reads them)
constexpr int F_THF = F_P0L;                // flux-average flag of a face
constexpr int F_PTOT = F_P0R;               // point total pressure, both sides summed
constexpr int F_PAVG = F_MEL;               // its theta-aware face average
constexpr int F_FPS = F_PHL;                // point physical fluxes, both sides summed
constexpr int F_PFA = F_PHR;                // their theta-aware face average

__device__ __forceinline__ double refMax(double a, double b) { return a < b ? b : a; }
__device__ __forceinline__ double refMin(double a, double b) { return b < a ? b : a; }

// bertaNormalizeWenoStencil (:747-773): divides by the largest magnitude and
// returns it (NaN-aware as the reference: a NaN is kept, never zeroed).
template <int N>
__device__ double wenoNormalize(double (&v)[N])
{
    double scale = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) scale = refMax(scale, fabs(v[i]));
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

// One side of f_wenoz_pointwise_faceF (:782-878) on normalized values. The
// candidate values q_k and the undivided differences are formed in FP64; the
// smoothness indicators and nonlinear weights in S (S = double reproduces
// the reference operation for operation); the weighted sum in FP64.
template <typename S>
__device__ __forceinline__ double wenozSide(double m2, double m1, double c, double p1, double p2)
{
    const S eps = S(1.0e-40), d0 = S(1.0/16.0), d1 = S(5.0/8.0), d2 = S(5.0/16.0);
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
    const double a0 = double(d0*(S(1.0) + tau5/(beta0 + eps)));
    const double a1 = double(d1*(S(1.0) + tau5/(beta1 + eps)));
    const double a2 = double(d2*(S(1.0) + tau5/(beta2 + eps)));
    const double asum = a0 + a1 + a2;
    return (a0*q0 + a1*q1 + a2*q2)/asum;
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

// 1. Raw Jameson flag on box+4 (f_berta_jameson_theta_rawF, :346-407): the
// normalized second difference of rho, p and the native |B|^2/2 per direction,
// the largest of the three, the 2-norm over directions, 1 below eta_d.
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
    auto magp = [&](int i, int j, int k) {
        const double bx = P[I(F_WA+5, i, j, k)], by = P[I(F_WA+6, i, j, k)], bz = P[I(F_WA+7, i, j, k)];
        return 0.5*(bx*bx + by*by + bz*bz); };
    double et[3];
#pragma unroll
    for (int dd = 0; dd < 3; ++dd)
    {
        const int m[3] = {x[0]-(dd == 0), x[1]-(dd == 1), x[2]-(dd == 2)};
        const int p[3] = {x[0]+(dd == 0), x[1]+(dd == 1), x[2]+(dd == 2)};
        const double er = eta1(P[I(F_WA, m[0], m[1], m[2])], P[I(F_WA, x[0], x[1], x[2])], P[I(F_WA, p[0], p[1], p[2])]);
        const double ep = eta1(P[I(F_WA+4, m[0], m[1], m[2])], P[I(F_WA+4, x[0], x[1], x[2])],
                               P[I(F_WA+4, p[0], p[1], p[2])]);
        const double em = eta1(magp(m[0], m[1], m[2]), magp(x[0], x[1], x[2]), magp(p[0], p[1], p[2]));
        et[dd] = refMax(er, refMax(ep, em));
    }
    const double eta = sqrt(et[0]*et[0] + et[1]*et[1] + et[2]*et[2]);
    P[I(F_THR, x[0], x[1], x[2])] = eta < v.etaD ? 1.0 : 0.0;
}

// 2. Dilated flag on box+5: the minimum over the cell and its six neighbours
// on box+3 (f_berta_dilate_thetaF, :410-434), 0 in the two outer layers (:500-503).
__global__ void __launch_bounds__(kB) k_b_theta(O4View v, double* __restrict__ S)
{
    const BoxDesc& b = blockBoxDesc(v);
    int L[3], H[3], x[3];
    for (int d = 0; d < 3; ++d) { L[d] = b.lo[d]-kG; H[d] = b.hi[d]+kG; }
    if (!regionPoint(b, L, H, boxThread(v, b), x)) return;
    const BoxIdx I(b);
    double* const P = S + b.sOff;
    bool inner = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) inner = inner && x[d] >= b.lo[d]-3 && x[d] <= b.hi[d]+3;
    double t = 0.0;
    if (inner)
    {
        t = P[I(F_THR, x[0], x[1], x[2])];
        t = refMin(t, P[I(F_THR, x[0]-1, x[1], x[2])]); t = refMin(t, P[I(F_THR, x[0]+1, x[1], x[2])]);
        t = refMin(t, P[I(F_THR, x[0], x[1]-1, x[2])]); t = refMin(t, P[I(F_THR, x[0], x[1]+1, x[2])]);
        t = refMin(t, P[I(F_THR, x[0], x[1], x[2]-1)]); t = refMin(t, P[I(F_THR, x[0], x[1], x[2]+1)]);
    }
    P[I(F_THETA, x[0], x[1], x[2])] = t;
}

// 3. Point primitives on box+4: UPoint = <U> - theta Lap<U>/24, V(UPoint)
// with the pressure clipped at 0 (:696-741, :212-275).
__global__ void __launch_bounds__(kB) k_b_wp(O4View v, const double* __restrict__ U, double* __restrict__ S)
{
    const BoxDesc& b = blockBoxDesc(v);
    int L[3], H[3], x[3];
    for (int d = 0; d < 3; ++d) { L[d] = b.lo[d]-4; H[d] = b.hi[d]+4; }
    if (!regionPoint(b, L, H, boxThread(v, b), x)) return;
    const BoxIdx I(b);
    double* const P = S + b.sOff;
    auto ui = [&](int i, int j, int k) {
        return b.uOff + (k-b.lo[2]+kG)*b.us[1] + (j-b.lo[1]+kG)*b.us[0] + (i-b.lo[0]+kG); };
    const bool smooth = P[I(F_THETA, x[0], x[1], x[2])] > 0.5;
    double up[8];
#pragma unroll
    for (int c = 0; c < 8; ++c)
    {
        const long long o = c*b.us[2];
        const double uc = U[ui(x[0],x[1],x[2])+o];
        const double lap = U[ui(x[0]+1,x[1],x[2])+o]+U[ui(x[0]-1,x[1],x[2])+o]+U[ui(x[0],x[1]+1,x[2])+o]
                          +U[ui(x[0],x[1]-1,x[2])+o]+U[ui(x[0],x[1],x[2]+1)+o]+U[ui(x[0],x[1],x[2]-1)+o]
                          -6.0*uc;
        up[c] = smooth ? uc - lap/24.0 : uc;
    }
    double w[8];
    sphToPrimPoint(w, up);
    double ke, me, vb;
    energiesNative(v.cellA + 9*angIdx(b, x[1], x[2]), w, ke, me, vb);
    w[4] = (w[4]-ke-me)*(v.gamma-1.0);
    if (w[4] < 0.0) w[4] = 0.0;
#pragma unroll
    for (int c = 0; c < 8; ++c) P[I(F_WPT+c, x[0], x[1], x[2])] = w[c];
}

// 4. Face points of direction D on the faces [lo_d, hi_d+1] x transverse
// [lo-2, hi+2]: WENOZ where the six-cell flag minimum is 1, else WENO3 (or PLM),
// floors, the Roe repair, then the point fluxes and total pressures of both
// sides summed. Reconstruction arithmetic in S (the normalization and the
// rescaling stay FP64).
template <int D, typename S_, bool ROE>
__global__ void __launch_bounds__(kB) k_b_faces(O4View v, double* __restrict__ S)
{
    const BoxDesc& b = blockBoxDesc(v);
    int L[3], H[3], x[3];
    for (int d = 0; d < 3; ++d) { L[d] = b.lo[d]-2; H[d] = b.hi[d]+2; }
    // the Powell face products also take the faces lo_d-1 and hi_d+2
    L[D] = b.lo[D] - (v.powell ? 1 : 0); H[D] = b.hi[D] + (v.powell ? 2 : 1);
    if (!regionPoint(b, L, H, boxThread(v, b), x)) return;
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
#pragma unroll 1
    for (int c = 0; c < 8; ++c)
    {
        if (tr > 0.5)
        {
            double s6[6];
#pragma unroll
            for (int o = -3; o <= 2; ++o) s6[o+3] = at(F_WPT+c, o);
            const double scale = wenoNormalize(s6);
            if (scale == 0.0) { l[c] = 0.0; r[c] = 0.0; }
            else
            {
                l[c] = scale*wenozSide<S_>(s6[0], s6[1], s6[2], s6[3], s6[4]);
                r[c] = scale*wenozSide<S_>(s6[5], s6[4], s6[3], s6[2], s6[1]);
            }
        }
        else
        {
            double s4[4];
#pragma unroll
            for (int o = -2; o <= 1; ++o) s4[o+2] = at(F_WPT+c, o);
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
        double cl[8], cr[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) { cl[c] = at(F_WA+c, -1); cr[c] = at(F_WA+c, 0); }
        if (usablePrimitive(cl) && usablePrimitive(cr))
        {
#pragma unroll
            for (int c = 0; c < 8; ++c) { l[c] = cl[c]; r[c] = cr[c]; }
            thF = 0.0;
        }
    }
    const double* A = v.faceA[D] + 9*angIdx(b, x[1], x[2]);
    double keL, meL, vbL, keR, meR, vbR, fl[8], fr[8];
    energiesNative(A, l, keL, meL, vbL);
    energiesNative(A, r, keR, meR, vbR);
    primFlux(fl, l, keL, meL, vbL, v.gamma, D);
    primFlux(fr, r, keR, meR, vbR, v.gamma, D);
#pragma unroll
    for (int c = 0; c < 8; ++c)
    {
        P[I(F_BWL+c, x[0], x[1], x[2])] = l[c];
        P[I(F_BWR+c, x[0], x[1], x[2])] = r[c];
        P[I(F_FPS+c, x[0], x[1], x[2])] = fl[c] + fr[c];
    }
    P[I(F_THF, x[0], x[1], x[2])] = thF;
    P[I(F_PTOT, x[0], x[1], x[2])] = (l[4] + meL) + (r[4] + meR);
    if constexpr (!ROE)
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

// 5. Theta-aware face averages of the summed point flux and total pressure
// on the faces [lo_d, hi_d+1] x transverse [lo-1, hi+1] (:1246-1266).
template <int D>
__global__ void __launch_bounds__(kB) k_b_avg(O4View v, double* __restrict__ S)
{
    const BoxDesc& b = blockBoxDesc(v);
    int L[3], H[3], x[3];
    for (int d = 0; d < 3; ++d) { L[d] = b.lo[d]-1; H[d] = b.hi[d]+1; }
    L[D] = b.lo[D]; H[D] = b.hi[D]+1;
    if (!regionPoint(b, L, H, boxThread(v, b), x)) return;
    const BoxIdx I(b);
    double* const P = S + b.sOff;
    constexpr int t1 = D == 0 ? 1 : 0, t2 = D == 2 ? 1 : 2;
    const bool smooth = P[I(F_THF, x[0], x[1], x[2])] > 0.5;
    auto avg = [&](int f) {
        const double c0 = P[I(f, x[0], x[1], x[2])];
        double lap = 0;
#pragma unroll
        for (int t = 0; t < 2; ++t)
        {
            const int tt = t == 0 ? t1 : t2;
            int y[3] = {x[0], x[1], x[2]}, z[3] = {x[0], x[1], x[2]};
            y[tt] += 1; z[tt] -= 1;
            lap += P[I(f, y[0], y[1], y[2])] - 2.0*c0 + P[I(f, z[0], z[1], z[2])];
        }
        return smooth ? c0 + lap/24.0 : c0; };
#pragma unroll
    for (int c = 0; c < 8; ++c) P[I(F_PFA+c, x[0], x[1], x[2])] = avg(F_FPS+c);
    P[I(F_PAVG, x[0], x[1], x[2])] = avg(F_PTOT);
}

// 6. Mapped products of the averaged flux at a radial face (bertaMapHostBox,
// :35-98): per matrix column one term (transverse corrections, then the
// product), the terms summed over columns. The radial flux kernel evaluates
// them at its own face; the angular directions use FaceOps::mapped.
__device__ __forceinline__ void bertaRadialMapped(const FaceOps<0>& op, const int x[3], double bmp[8])
{
    constexpr int t1 = 1, t2 = 2;
#pragma unroll
    for (int s = 0; s < 2; ++s)
    {
        const int c = s == 0 ? 0 : 4;
        double term = 0;
        term += op.Dm(0, t1, x)*op.Df(F_PFA+c, t1, x)*(1.0/12.0);
        term += op.Dm(0, t2, x)*op.Df(F_PFA+c, t2, x)*(1.0/12.0);
        term += op.M(0, x)*op.Fv(F_PFA+c, x)*1.0;
        double sum = 0;
        sum += term;
        bmp[c] = sum;
    }
#pragma unroll
    for (int grp = 0; grp < 2; ++grp)
    {
        const int c0 = grp == 0 ? 1 : 5;
#pragma unroll
        for (int r = 0; r < 3; ++r)
        {
            double sum = 0;
#pragma unroll
            for (int col = 0; col < 3; ++col)
            {
                double term = 0;
                term += op.Dm(1+3*r+col, t1, x)*op.Df(F_PFA+c0+col, t1, x)*(1.0/12.0);
                term += op.Dm(1+3*r+col, t2, x)*op.Df(F_PFA+c0+col, t2, x)*(1.0/12.0);
                term += op.M(1+3*r+col, x)*op.Fv(F_PFA+c0+col, x)*1.0;
                sum += term;
            }
            bmp[c0+r] = sum;
        }
    }
}

// 7. Face flux on [lo_d, hi_d+1] x [lo, hi]: radial-area tensor product of the
// mapped products, pressure force ntrow <(p_tot averaged) adjDr>, the
// theta-aware averaged Roe dissipation with its metric products, and
// F = (Sum - D)/2. Evaluated one input component at a time (its stencil of
// averaged fluxes and dissipation, then its contributions to the outputs it
// enters), which bounds the live registers; every output accumulates its
// terms in the same order as the term-by-term form.
template <int D, bool ROE>
__global__ void __launch_bounds__(kB) k_b_flux(O4View v, double* __restrict__ S)
{
    const BoxDesc& b = blockBoxDesc(v);
    int L[3], H[3], x[3];
    for (int d = 0; d < 3; ++d) { L[d] = b.lo[d]; H[d] = b.hi[d]; }
    H[D] = b.hi[D]+1;
    if (!regionPoint(b, L, H, boxThread(v, b), x)) return;
    const BoxIdx I(b);
    double* const P = S + b.sOff;
    constexpr int t1 = D == 0 ? 1 : 0, t2 = D == 2 ? 1 : 2;
    const FaceOps<D> op(v, b, I, P);
    auto Fv = [&](int f, const int y[3]) { return op.Fv(f, y); };
    const double adjX = areaOf(v, b, D, x[0]);
    auto dAdj = [&](int i) { return 0.5*areaOf(v, b, D, i+1) - 0.5*areaOf(v, b, D, i-1); };
    const double dA = D == 0 ? 0.0 : dAdj(x[0]);
    auto shift = [&](int f, int t, int o) { int y[3] = {x[0], x[1], x[2]}; y[t] += o; return Fv(f, y); };
    auto shift2 = [&](int f, int o0, int o2) { int y[3] = {x[0]+o0, x[1], x[2]}; y[t2] += o2; return Fv(f, y); };
    // Df(f, t) at x and, for the angular radial difference, the half radial
    // differences Gv(f, .) at x and x +- e_t2 (Gv = (F(+e_r) - F(-e_r))/2)
    auto Df = [&](int f, int t) { return 0.5*shift(f, t, 1) - 0.5*shift(f, t, -1); };
    auto Gv0 = [&](int f) { return 0.5*shift(f, 0, 1) - 0.5*shift(f, 0, -1); };
    auto DGv = [&](int f) {
        const double gp = 0.5*shift2(f, 1, 1) - 0.5*shift2(f, -1, 1);
        const double gm = 0.5*shift2(f, 1, -1) - 0.5*shift2(f, -1, -1);
        return 0.5*gp - 0.5*gm; };
    // face-average Roe dissipation of component f (theta-aware)
    const bool smooth = Fv(F_THF, x) > 0.5;
    auto Davg = [&](int f) {
        const double c0 = Fv(f, x);
        double lap = 0;
#pragma unroll
        for (int t = 0; t < 2; ++t)
        {
            const int tt = t == 0 ? t1 : t2;
            lap += shift(f, tt, 1) - 2.0*c0 + shift(f, tt, -1);
        }
        return smooth ? c0 + lap/24.0 : c0; };
    // Fsum of a component from its mapped products (radial: bertaRadialMapped's
    // per-column terms; angular: FaceOps::mapped and mappedRadialDiff)
    auto fsum = [&](double mp, double dmp) {
        double val = 0;
        if (D != 0) val += dmp*dA*(1.0/12.0);
        val += mp*adjX*1.0;
        return val; };
    // <p_tot adjDr> at y: radial tensor product of the averaged total pressure
    auto tmp3 = [&](const int y[3]) {
        double val = 0;
        if (D != 0)
        {
            const int a[3] = {y[0]+1, y[1], y[2]}, z[3] = {y[0]-1, y[1], y[2]};
            const double dv = 0.5*Fv(F_PAVG, a) - 0.5*Fv(F_PAVG, z);
            val += dv*dAdj(y[0])*(1.0/12.0);
        }
        val += Fv(F_PAVG, y)*areaOf(v, b, D, y[0])*1.0;
        return val; };
    double press[3];
    {
        const double t0v = tmp3(x);
        double dT[2] = {0.0, 0.0};
#pragma unroll
        for (int t = D == 0 ? 0 : 1; t < 2; ++t)
        {
            const int tt = t == 0 ? t1 : t2;
            int y[3] = {x[0], x[1], x[2]}, z[3] = {x[0], x[1], x[2]};
            y[tt] += 1; z[tt] -= 1;
            dT[t] = 0.5*tmp3(y) - 0.5*tmp3(z);
        }
#pragma unroll
        for (int r = 0; r < 3; ++r)
        {
            double val = 0;
            if (D == 0) val += op.Dm(10+r, t1, x)*dT[0]*(1.0/12.0);
            val += op.Dm(10+r, t2, x)*dT[1]*(1.0/12.0);
            val += op.M(10+r, x)*t0v*1.0;
            press[r] = val;
        }
    }
    const double speed = ROE ? 0.0 : Fv(F_SPD, x);
    // scalar components 0 and 4
#pragma unroll 1
    for (int s = 0; s < 2; ++s)
    {
        const int c = s == 0 ? 0 : 4;
        double mp = 0, dmp = 0;
        if (D == 0)
        {
            double term = 0;
            term += op.Dm(0, t1, x)*Df(F_PFA+c, t1)*(1.0/12.0);
            term += op.Dm(0, t2, x)*Df(F_PFA+c, t2)*(1.0/12.0);
            term += op.M(0, x)*Fv(F_PFA+c, x)*1.0;
            mp += term;
        }
        else
        {
            mp += op.Dm(0, t2, x)*Df(F_PFA+c, t2)*(1.0/12.0);
            mp += op.M(0, x)*Fv(F_PFA+c, x);
            dmp += op.Dm(0, t2, x)*DGv(F_PFA+c)*(1.0/12.0);
            dmp += op.M(0, x)*Gv0(F_PFA+c);
        }
        const double fs = fsum(mp, dmp);
        const double da = Davg(F_DPT+c);
        double fd;
        if constexpr (ROE)
        {
            fd = 0;
            fd += op.Dm(0, t1, x)*Df(F_DPT+c, t1)*(1.0/12.0);
            fd += op.Dm(0, t2, x)*Df(F_DPT+c, t2)*(1.0/12.0);
            fd += op.M(0, x)*da;
        }
        else fd = (op.M(0, x)*da)*speed;
        P[I(F_FLUX+8*D+c, x[0], x[1], x[2])] = 0.5*(fs-fd);
    }
    // vector components: momentum (grp 0, with the pressure force) and field
#pragma unroll 1
    for (int grp = 0; grp < 2; ++grp)
    {
        const int c0 = grp == 0 ? 1 : 5;
        double mp[3] = {0, 0, 0}, dmp[3] = {0, 0, 0}, fd[3] = {0, 0, 0};
#pragma unroll 1
        for (int col = 0; col < 3; ++col)
        {
            const int f = F_PFA+c0+col, g = F_DPT+c0+col;
            const double fx = Fv(f, x), df2 = Df(f, t2);
            const double df1 = D == 0 ? Df(f, t1) : 0.0;
            const double gx = D == 0 ? 0.0 : Gv0(f), dg = D == 0 ? 0.0 : DGv(f);
            const double da = Davg(g);
            const double dd1 = ROE ? Df(g, t1) : 0.0, dd2 = ROE ? Df(g, t2) : 0.0;
#pragma unroll
            for (int r = 0; r < 3; ++r)
            {
                const int m = 1+3*r+col;
                if (D == 0)
                {
                    double term = 0;
                    term += op.Dm(m, t1, x)*df1*(1.0/12.0);
                    term += op.Dm(m, t2, x)*df2*(1.0/12.0);
                    term += op.M(m, x)*fx*1.0;
                    mp[r] += term;
                }
                else
                {
                    mp[r] += op.Dm(m, t2, x)*df2*(1.0/12.0);
                    mp[r] += op.M(m, x)*fx;
                    dmp[r] += op.Dm(m, t2, x)*dg*(1.0/12.0);
                    dmp[r] += op.M(m, x)*gx;
                }
                if constexpr (ROE)
                {
                    fd[r] += op.Dm(m, t1, x)*dd1*(1.0/12.0);
                    fd[r] += op.Dm(m, t2, x)*dd2*(1.0/12.0);
                    fd[r] += op.M(m, x)*da;
                }
                else fd[r] += op.M(m, x)*da;
            }
        }
#pragma unroll
        for (int r = 0; r < 3; ++r)
        {
            double fs = fsum(mp[r], dmp[r]);
            if (grp == 0) fs = fs + press[r];
            const double d = ROE ? fd[r] : fd[r]*speed;
            P[I(F_FLUX+8*D+c0+r, x[0], x[1], x[2])] = 0.5*(fs-d);
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
    if (v.roe) hipLaunchKernelGGL((k_b_faces<D, S_, true>), grid, dim3(kB), 0, s, v, scratch);
    else hipLaunchKernelGGL((k_b_faces<D, S_, false>), grid, dim3(kB), 0, s, v, scratch);
    if (v.powell) hipLaunchKernelGGL(k_b_pw_face<D>, grid, dim3(kB), 0, s, v, U, scratch);
    hipLaunchKernelGGL(k_b_avg<D>, grid, dim3(kB), 0, s, v, scratch);
    if (v.roe) hipLaunchKernelGGL((k_o4_roe<D, S_>), grid, dim3(kB), 0, s, v, scratch);
    if (v.roe) hipLaunchKernelGGL((k_b_flux<D, true>), grid, dim3(kB), 0, s, v, scratch);
    else hipLaunchKernelGGL((k_b_flux<D, false>), grid, dim3(kB), 0, s, v, scratch);
}

// Cell stage and the three direction passes of the Berta operator.
template <typename S_>
void bertaFluxes(const O4View& v, dim3 grid, hipStream_t s, const double* U, double* scratch)
{
    hipLaunchKernelGGL(k_o4_wbar, grid, dim3(kB), 0, s, v, U, scratch);
    hipLaunchKernelGGL(k_b_thraw, grid, dim3(kB), 0, s, v, scratch);
    hipLaunchKernelGGL(k_b_theta, grid, dim3(kB), 0, s, v, scratch);
    hipLaunchKernelGGL(k_b_wp, grid, dim3(kB), 0, s, v, U, scratch);
    bertaDirectionPass<0, S_>(v, grid, s, U, scratch);
    bertaDirectionPass<1, S_>(v, grid, s, U, scratch);
    bertaDirectionPass<2, S_>(v, grid, s, U, scratch);
}
