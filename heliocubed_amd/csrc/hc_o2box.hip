// Second-order per-box operator of the HelioCubed MHD RHS on MI355X:
// secondOrderOperator (src/BoxOp_SecondOrder.H:155-370; src/ = HelioCubed
// reference repository) as StageLevel calls it at order 2 with Cartesian
// primitives (src/BoxOp_EulerCubedSphereImplem.H:1003-1004), over a table of
// boxes evaluated in the same launches: PLM face states of the box's ghosted
// primitives (reconstructSecondOrderFace), the frame flux through the face's
// integrated normal (secondOrderFaceFlux, src/MHD_SecondOrderFlux.H:34-74),
// the flux divergence accumulated into the box RHS direction by direction, and
// with divergence_cleaning 1 the Powell source of the face magnetic fluxes.
// Without shared panel states every face, seam faces included, takes the
// states of the box's own ghost cells (the caller's exchange fills them).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
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
constexpr int kB2 = 256;
constexpr int kG2 = 2;   // ghost width of the second-order operator

struct Box2
{
    int p, lo[3], n[3];
    long long wOff, rOff, fOff[3];   // element offsets of W (box grown by 2), RHS, face arrays
    long long blk[4];                // first block of the box in the launches of faces 0..2 and cells
};

struct O2View
{
    int nr, na, limiter, roe, powell;
    double gamma, entropy;
    const Box2* boxes;
    const int* blockBox[4];          // box of each block of the face launches 0..2 and the cell launch
    const double* frame[3];          // [6][fk][fj][10]: frame rows + |normal|
    const double* normal[3];         // [6][fk][fj][3] integrated face normals (Powell)
    const double* radial;            // [nr+1][2]: area0, areaT of radial index i = 0 .. nr
};

template <int D>
__device__ __forceinline__ long long faceCount(const Box2& b)
{
    return (long long)(b.n[0]+(D == 0))*(b.n[1]+(D == 1))*(b.n[2]+(D == 2));
}

// Face fluxes of direction D of every box; magnetic face fluxes when the
// Powell source is on.
template <int D, int RS>
__global__ void __launch_bounds__(kB2) k_o2b_face(O2View v, const double* __restrict__ W, double* __restrict__ F,
                                                  double* __restrict__ BF)
{
    const Box2& b = v.boxes[v.blockBox[D][blockIdx.x]];
    const long long t = (blockIdx.x - b.blk[D])*(long long)kB2 + threadIdx.x;
    const int m0 = b.n[0]+(D == 0), m1 = b.n[1]+(D == 1);
    const long long nf = faceCount<D>(b);
    if (t >= nf) return;
    const int a0 = (int)(t % m0), a1 = (int)((t / m0) % m1), a2 = (int)(t / ((long long)m0*m1));
    const int i = b.lo[0]+a0, j = b.lo[1]+a1, k = b.lo[2]+a2;
    const int e0 = b.n[0]+2*kG2, e1 = b.n[1]+2*kG2;
    const long long wcs = (long long)e0*e1*(b.n[2]+2*kG2);
    auto load = [&](int o, double out[8]) {
        const long long x = b.wOff + ((long long)(a2+kG2+(D == 2)*o)*e1 + (a1+kG2+(D == 1)*o))*e0 + (a0+kG2+(D == 0)*o);
#pragma unroll
        for (int c = 0; c < 8; ++c) out[c] = W[x + c*wcs];
    };
    double mm[8], m[8], c[8], pp[8], l[8], r[8];
    load(-2, mm); load(-1, m); load(0, c); load(1, pp);
    hc::plmFaceStates(l, r, mm, m, c, pp, v.limiter);
    const int fj = v.na + (D == 1), fk = v.na + (D == 2);
    const long long col = ((long long)b.p*fk + k)*fj + j;
    const double* fr = v.frame[D] + 10*col;
    const double rarea = v.radial[2*i + (D == 0 ? 0 : 1)];
    double out[8];
    hc::o2FaceFluxMixed<double, RS>(out, l, r, fr, rarea*fr[9], v.gamma, v.entropy, RS == 2);
#pragma unroll
    for (int q = 0; q < 8; ++q) F[b.fOff[D] + q*nf + t] = out[q];
    if (v.powell)
    {
        const double* nrm = v.normal[D] + 3*col;
        double f = 0;
#pragma unroll
        for (int x = 0; x < 3; ++x) f += 0.5*(l[5+x]+r[5+x])*nrm[x]*rarea;
        BF[b.fOff[D]/8 + t] = f;
    }
}

// RHS of every valid cell: acc = RHS; acc += (-scale/dxi_d)(F_d(high) -
// F_d(low)) for d = 0, 1, 2; then the Powell source of -div B of the magnetic
// face fluxes times (B/4pi, v, v.B/4pi) of the cell's primitives.
__global__ void __launch_bounds__(kB2) k_o2b_div(O2View v, const double* __restrict__ W, const double* __restrict__ F0,
                                                 const double* __restrict__ F1, const double* __restrict__ F2,
                                                 const double* __restrict__ B0, const double* __restrict__ B1,
                                                 const double* __restrict__ B2, double* __restrict__ R, double scale)
{
    const Box2& b = v.boxes[v.blockBox[3][blockIdx.x]];
    const long long t = (blockIdx.x - b.blk[3])*(long long)kB2 + threadIdx.x;
    const long long nc = (long long)b.n[0]*b.n[1]*b.n[2];
    if (t >= nc) return;
    const int a0 = (int)(t % b.n[0]), a1 = (int)((t / b.n[0]) % b.n[1]), a2 = (int)(t / ((long long)b.n[0]*b.n[1]));
    const double* F[3] = {F0, F1, F2};
    const double* BF[3] = {B0, B1, B2};
    const double dxi[3] = {1.0/v.nr, 1.0/v.na, 1.0/v.na};
    long long lo[3], hi[3], nf[3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        const int m0 = b.n[0]+(d == 0), m1 = b.n[1]+(d == 1), m2 = b.n[2]+(d == 2);
        nf[d] = (long long)m0*m1*m2;
        lo[d] = ((long long)a2*m1 + a1)*m0 + a0;
        hi[d] = lo[d] + (d == 0 ? 1 : d == 1 ? m0 : (long long)m0*m1);
    }
    double acc[8];
#pragma unroll
    for (int q = 0; q < 8; ++q)
    {
        acc[q] = R[b.rOff + q*nc + t];
#pragma unroll
        for (int d = 0; d < 3; ++d)
        {
            const double* f = F[d] + b.fOff[d] + q*nf[d];
            acc[q] += (-scale/dxi[d])*(f[hi[d]]-f[lo[d]]);
        }
    }
    if (v.powell)
    {
        double md = 0;
#pragma unroll
        for (int d = 0; d < 3; ++d)
        {
            const double* f = BF[d] + b.fOff[d]/8;
            md += (-scale/dxi[d])*(f[hi[d]]-f[lo[d]]);
        }
        const int e0 = b.n[0]+2*kG2, e1 = b.n[1]+2*kG2;
        const long long wcs = (long long)e0*e1*(b.n[2]+2*kG2);
        const long long x = b.wOff + ((long long)(a2+kG2)*e1 + (a1+kG2))*e0 + (a0+kG2);
        double vb = 0;
#pragma unroll
        for (int d = 0; d < 3; ++d)
        {
            const double bb = W[x + (5+d)*wcs], vv = W[x + (1+d)*wcs];
            acc[1+d] += md*bb/(4.0*hc::kPi);
            acc[5+d] += md*vv;
            vb += vv*bb;
        }
        acc[4] += md*vb/(4.0*hc::kPi);
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) R[b.rOff + q*nc + t] = acc[q];
}

template <typename T>
T* upload2(const std::vector<T>& h, std::vector<void*>& owned)
{
    T* p = nullptr;
    HC_THROW(hipMalloc(&p, std::max<size_t>(1, h.size())*sizeof(T)));
    owned.push_back(p);
    if (!h.empty()) HC_THROW(hipMemcpy(p, h.data(), h.size()*sizeof(T), hipMemcpyHostToDevice));
    return p;
}
} // namespace

struct hc_boxop2
{
    hc_options opt{};
    O2View view{};
    int nbox = 0;
    long long wCount = 0, rCount = 0, fCount[3] = {0, 0, 0}, blocks[4] = {0, 0, 0, 0};
    double* scratchF[3] = {nullptr, nullptr, nullptr};   // face fluxes when the caller passes none
    double* scratchB[3] = {nullptr, nullptr, nullptr};   // magnetic face fluxes (Powell)
    std::vector<void*> owned;
    ~hc_boxop2() { for (void* p : owned) (void)hipFree(p); }
};

extern "C" {

int hc_boxop2_create(const hc_shell* shell, const hc_options* options, const hc_box* boxes, int nbox,
                     hc_boxop2** out)
{
    if (!shell || !options || !boxes || !out || nbox <= 0) return hc::fail("hc_boxop2_create: bad argument");
    *out = nullptr;
    if (options->spatial_order != 2) return hc::fail("hc_boxop2_create: spatial_order must be 2");
    if (options->riemann_solver != 1 && options->riemann_solver != 2)
        return hc::fail("hc_boxop2_create: riemann_solver must be 1 (Rusanov) or 2 (Roe)");
    if (options->limiter < 0 || options->limiter > 3) return hc::fail("hc_boxop2_create: limiter must be 0..3");
    if (options->precision != 0) return hc::fail("hc_boxop2_create: the per-box second-order operator is FP64 (precision 0)");
    if (options->spatial_operator == 2) return hc::fail("hc_boxop2_create: the Berta operator is fourth order");
    if (options->divergence_cleaning != 0 && options->divergence_cleaning != 1)
        return hc::fail("hc_boxop2_create: divergence_cleaning must be 0 or 1");
    if (shell->n_rad < 1 || shell->n_ang < 2 || !(shell->outer_radius > shell->inner_radius && shell->inner_radius > 0))
        return hc::fail("hc_boxop2_create: bad shell");
    for (int n = 0; n < nbox; ++n)
    {
        const hc_box& b = boxes[n];
        if (b.block < 0 || b.block > 5) return hc::fail("hc_boxop2_create: box block must be 0..5");
        const int lim[3] = {shell->n_rad, shell->n_ang, shell->n_ang};
        for (int d = 0; d < 3; ++d)
            if (b.lo[d] < 0 || b.hi[d] < b.lo[d] || b.hi[d] >= lim[d])
                return hc::fail("hc_boxop2_create: box outside its panel");
    }
    try
    {
        auto* B = new hc_boxop2();
        std::unique_ptr<hc_boxop2> guard(B);
        B->opt = *options;
        O2View& v = B->view;
        const int nr = shell->n_rad, na = shell->n_ang;
        v.nr = nr; v.na = na; v.limiter = options->limiter; v.roe = options->riemann_solver == 2;
        v.powell = options->divergence_cleaning == 1;
        v.gamma = options->gamma; v.entropy = options->entropy_fix;
        const hc::ShellDims dims{nr, na, kG2, shell->inner_radius, shell->outer_radius, shell->radial_stretch};
        std::vector<double> radial(2*(nr+1));
        for (int i = 0; i <= nr; ++i)
        {
            const hc::RadialCell c = hc::radialCell(dims, i);
            radial[2*i] = c.area0; radial[2*i+1] = c.areaT;
        }
        v.radial = upload2(radial, B->owned);
        for (int d = 0; d < 3; ++d)
        {
            const int fj = na + (d == 1), fk = na + (d == 2);
            std::vector<double> fr(10ull*6*fj*fk), nrm(3ull*6*fj*fk);
            for (int p = 0; p < 6; ++p) for (int k = 0; k < fk; ++k) for (int j = 0; j < fj; ++j)
            {
                const hc::Vec3 nv = hc::faceNormal(na, p, d, j, k);
                const hc::FaceFrame F = hc::makeFaceFrame(nv.x);
                double* o = &fr[10*(((size_t)p*fk+k)*fj+j)];
                for (int a = 0; a < 3; ++a) for (int c = 0; c < 3; ++c) o[3*a+c] = F.f[a][c];
                o[9] = F.norm;
                for (int a = 0; a < 3; ++a) nrm[3*(((size_t)p*fk+k)*fj+j)+a] = nv.x[a];
            }
            v.frame[d] = upload2(fr, B->owned);
            v.normal[d] = v.powell ? upload2(nrm, B->owned) : nullptr;
        }
        std::vector<Box2> bx;
        std::vector<int> blockBox[4];
        long long w = 0, r = 0, f[3] = {0, 0, 0};
        for (int n = 0; n < nbox; ++n)
        {
            Box2 b{};
            b.p = boxes[n].block;
            for (int d = 0; d < 3; ++d) { b.lo[d] = boxes[n].lo[d]; b.n[d] = boxes[n].hi[d]-boxes[n].lo[d]+1; }
            b.wOff = w; b.rOff = r;
            w += 8LL*(b.n[0]+2*kG2)*(b.n[1]+2*kG2)*(b.n[2]+2*kG2);
            const long long nc = (long long)b.n[0]*b.n[1]*b.n[2];
            r += 8*nc;
            for (int d = 0; d < 3; ++d)
            {
                b.fOff[d] = f[d];
                f[d] += 8LL*(b.n[0]+(d == 0))*(b.n[1]+(d == 1))*(b.n[2]+(d == 2));
            }
            for (int q = 0; q < 4; ++q)
            {
                const long long cnt = q < 3 ? (long long)(b.n[0]+(q == 0))*(b.n[1]+(q == 1))*(b.n[2]+(q == 2)) : nc;
                b.blk[q] = (long long)blockBox[q].size();
                for (long long s = 0; s < (cnt + kB2 - 1)/kB2; ++s) blockBox[q].push_back(n);
            }
            bx.push_back(b);
        }
        for (int q = 0; q < 4; ++q)
        {
            if (blockBox[q].size() > 0x7fffffffULL) throw std::runtime_error("box table too large for one launch");
            B->blocks[q] = (long long)blockBox[q].size();
            v.blockBox[q] = upload2(blockBox[q], B->owned);
        }
        v.boxes = upload2(bx, B->owned);
        B->wCount = w; B->rCount = r;
        for (int d = 0; d < 3; ++d)
        {
            B->fCount[d] = f[d];
            HC_THROW(hipMalloc(&B->scratchF[d], f[d]*sizeof(double)));
            B->owned.push_back(B->scratchF[d]);
            if (v.powell)
            {
                HC_THROW(hipMalloc(&B->scratchB[d], std::max<long long>(1, f[d]/8)*sizeof(double)));
                B->owned.push_back(B->scratchB[d]);
            }
        }
        B->nbox = nbox;
        *out = guard.release();
        return 0;
    }
    catch (const std::exception& e) { return hc::fail(std::string("hc_boxop2_create: ") + e.what()); }
}

void hc_boxop2_destroy(hc_boxop2* op) { delete op; }

long long hc_boxop2_count(const hc_boxop2* op, int what)
{
    if (!op) return -1;
    if (what == 0) return op->wCount;
    if (what == 1) return op->rCount;
    if (what >= 2 && what <= 4) return op->fCount[what-2];
    return -1;
}

int hc_boxop2_apply(hc_boxop2* op, const double* W, double* rhs, double* F0, double* F1, double* F2, double scale,
                    hc_stream stream)
{
    if (!op || !W || !rhs) return hc::fail("hc_boxop2_apply: null argument");
    if (!std::isfinite(scale)) return hc::fail("hc_boxop2_apply: the divergence scale must be finite");
    try
    {
        const O2View& v = op->view;
        const hipStream_t s = static_cast<hipStream_t>(stream);
        double* F[3] = {F0 ? F0 : op->scratchF[0], F1 ? F1 : op->scratchF[1], F2 ? F2 : op->scratchF[2]};
        if (v.roe)
        {
            hipLaunchKernelGGL((k_o2b_face<0, 2>), dim3((unsigned)op->blocks[0]), dim3(kB2), 0, s, v, W, F[0], op->scratchB[0]);
            hipLaunchKernelGGL((k_o2b_face<1, 2>), dim3((unsigned)op->blocks[1]), dim3(kB2), 0, s, v, W, F[1], op->scratchB[1]);
            hipLaunchKernelGGL((k_o2b_face<2, 2>), dim3((unsigned)op->blocks[2]), dim3(kB2), 0, s, v, W, F[2], op->scratchB[2]);
        }
        else
        {
            hipLaunchKernelGGL((k_o2b_face<0, 1>), dim3((unsigned)op->blocks[0]), dim3(kB2), 0, s, v, W, F[0], op->scratchB[0]);
            hipLaunchKernelGGL((k_o2b_face<1, 1>), dim3((unsigned)op->blocks[1]), dim3(kB2), 0, s, v, W, F[1], op->scratchB[1]);
            hipLaunchKernelGGL((k_o2b_face<2, 1>), dim3((unsigned)op->blocks[2]), dim3(kB2), 0, s, v, W, F[2], op->scratchB[2]);
        }
        hipLaunchKernelGGL(k_o2b_div, dim3((unsigned)op->blocks[3]), dim3(kB2), 0, s, v, W, F[0], F[1], F[2],
                           op->scratchB[0], op->scratchB[1], op->scratchB[2], rhs, scale);
        HC_THROW(hipGetLastError());
        return 0;
    }
    catch (const std::exception& e) { return hc::fail(std::string("hc_boxop2_apply: ") + e.what()); }
}

} // extern "C"
