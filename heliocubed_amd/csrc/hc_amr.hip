// AMR coarse-to-fine transfers of one patch (the data the AMR driver moves
// between levels before each stage):
//
//  * hc_amr_interp4: fourth-order averages of the fine cells from the coarse
//    cell averages, amrInterpolateFourthOrder (src/MHD_AMRInterpolation.H:39-161).
//    Along each refined direction the five coarse averages around the parent
//    define the quartic with those averages; its average over the child cell
//    is a fixed five-weight combination (the weights depend on the ratio and
//    the child index only). The tensor product is applied as three
//    one-dimensional passes, radial first, as the reference does.
//  * hc_amr_ghosts2: second-order fine ghost cells, src/MHD_AMR.H:1282-1393:
//    the tensor quadratic average (q = x^2 + (1/r^2 - 1)/12, weights
//    ((q - x)/2, 1 - q, (q + x)/2)) of the coarse J*U and of J, their quotient,
//    a positivity check that falls back to the parent cell, then Cartesian
//    primitives.
//  * hc_amr_average_down: restriction of fine averages to their parents
//    (Stencil::AvgDown in restrictTo, src/MHD_AMR.H:596-606).
//  * hc_amr_prolong4: conservative fourth-order refinement of new fine cells
//    with the positivity-preserving bisection (prolongFourthOrder,
//    src/MHD_AMRFourthOrder.H:226-314).
//
// Patches use the Proto BoxData layout: component-major, index 0 fastest,
// inclusive lo..hi in the level's (possibly negative ghost) index space.
#include "hc_common.hpp"
#include "hc_geometry.hpp"
#include "hc_physics.hpp"

#include <cmath>
#include <mutex>
#include <utility>
#include <vector>

// Per-parent flags of the allocation-free conservative refinement
// (hc_amr_prolong4_noalloc) and, after the last flag slot, their count: a
// file-scope device array, so that entry point calls no allocator. Each
// device the library runs on holds its own copy (4 MiB) until the process
// ends; host code reaches it only through the address hipGetSymbolAddress
// returns on that device.
constexpr long long kNoAllocParents = 1LL << 20;
static __device__ __align__(256) int scratch[kNoAllocParents + 64];

namespace hc
{
namespace
{
constexpr int kAmrBlock = 256;
constexpr int kMaxRatio = 8;

struct Patch
{
    int lo[3], hi[3];
    __host__ __device__ long long n(int d) const { return hi[d] - lo[d] + 1; }
    __host__ __device__ long long count() const { return n(0)*n(1)*n(2); }
    __host__ __device__ bool contains(int a, int d) const { return a >= lo[d] && a <= hi[d]; }
    __host__ __device__ long long at(int c, int i, int j, int k) const
    { return ((c*n(2) + (k - lo[2]))*n(1) + (j - lo[1]))*n(0) + (i - lo[0]); }
};

__host__ __device__ inline int floorDiv(int p, int r) { return (int)floor(double(p)/r); }

struct Interp4View
{
    Patch src, dst;
    int ncomp;
    int ratio[3];
    double w[3][kMaxRatio][5];   // child weights per direction
};

__global__ void __launch_bounds__(kAmrBlock) k_amr_interp4(Interp4View v, const double* __restrict__ src,
                                                           double* __restrict__ dst)
{
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= v.dst.count()) return;
    int p[3];
    p[0] = v.dst.lo[0] + (int)(t % v.dst.n(0));
    p[1] = v.dst.lo[1] + (int)((t / v.dst.n(0)) % v.dst.n(1));
    p[2] = v.dst.lo[2] + (int)(t / (v.dst.n(0)*v.dst.n(1)));
    int q[3], rad[3];
    const double* w[3];
    bool inside = true;
    for (int d = 0; d < 3; ++d)
    {
        q[d] = floorDiv(p[d], v.ratio[d]);
        w[d] = v.w[d][p[d] - q[d]*v.ratio[d]];
        rad[d] = v.ratio[d] > 1 ? 2 : 0;
        for (int k = -rad[d]; k <= rad[d]; ++k)
            if (w[d][k+2] != 0.0 && !v.src.contains(q[d]+k, d)) inside = false;
    }
    for (int c = 0; c < v.ncomp; ++c)
    {
        double value = NAN;
        if (inside)
        {
            double s2 = 0.0;
            for (int k2 = -rad[2]; k2 <= rad[2]; ++k2)
            {
                if (w[2][k2+2] == 0.0) continue;
                double s1 = 0.0;
                for (int k1 = -rad[1]; k1 <= rad[1]; ++k1)
                {
                    if (w[1][k1+2] == 0.0) continue;
                    double s0 = 0.0;
                    for (int k0 = -rad[0]; k0 <= rad[0]; ++k0)
                    {
                        if (w[0][k0+2] == 0.0) continue;
                        s0 += w[0][k0+2]*src[v.src.at(c, q[0]+k0, q[1]+k1, q[2]+k2)];
                    }
                    s1 += w[1][k1+2]*s0;
                }
                s2 += w[2][k2+2]*s1;
            }
            value = s2;
        }
        dst[v.dst.at(c, p[0], p[1], p[2])] = value;
    }
}

struct Ghost2View
{
    Patch src, dst, owned;
    int ratio[3];
    double gamma;
};

__global__ void __launch_bounds__(kAmrBlock) k_amr_ghosts2(Ghost2View v, const double* __restrict__ src,
                                                           double* __restrict__ dst)
{
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= v.dst.count()) return;
    int p[3];
    p[0] = v.dst.lo[0] + (int)(t % v.dst.n(0));
    p[1] = v.dst.lo[1] + (int)((t / v.dst.n(0)) % v.dst.n(1));
    p[2] = v.dst.lo[2] + (int)(t / (v.dst.n(0)*v.dst.n(1)));
    if (v.owned.contains(p[0], 0) && v.owned.contains(p[1], 1) && v.owned.contains(p[2], 2)) return;
    int q[3];
    double w[3][3];
    for (int d = 0; d < 3; ++d)
    {
        const int r = v.ratio[d];
        q[d] = floorDiv(p[d], r);
        const double x = (p[d] + 0.5)/r - (q[d] + 0.5);
        const double quadratic = x*x + (1.0/(double(r)*r) - 1)/12;
        w[d][0] = 0.5*(quadratic - x);
        w[d][1] = 1 - quadratic;
        w[d][2] = 0.5*(quadratic + x);
        // every donor of the stencil must lie in the coarse patch
        const int reach = r == 1 ? 0 : 1;
        if (!v.src.contains(q[d]-reach, d) || !v.src.contains(q[d]+reach, d)) return;
    }
    if (!isfinite(src[v.src.at(8, q[0], q[1], q[2])])) return;
    double u[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double j = 0.0;
    for (int o2 = -1; o2 <= 1; ++o2)
        for (int o1 = -1; o1 <= 1; ++o1)
            for (int o0 = -1; o0 <= 1; ++o0)
            {
                const int off[3] = {o0, o1, o2};
                bool used = true;
                for (int d = 0; d < 3; ++d) if (v.ratio[d] == 1 && off[d]) used = false;
                if (!used) continue;
                double weight = 1;
                for (int d = 0; d < 3; ++d) weight *= w[d][off[d]+1];
                if (weight == 0) continue;
                const int a = q[0]+o0, b = q[1]+o1, c3 = q[2]+o2;
                j += weight*src[v.src.at(8, a, b, c3)];
                for (int c = 0; c < 8; ++c) u[c] += weight*src[v.src.at(c, a, b, c3)];
            }
    bool valid = j > 0 && isfinite(j);
    for (int c = 0; c < 8; ++c) { u[c] /= j; valid &= (bool)isfinite(u[c]); }
    double thermal = u[4];
    if (u[0] > 0)
        for (int d = 0; d < 3; ++d) thermal -= 0.5*u[1+d]*u[1+d]/u[0] + u[5+d]*u[5+d]/(8*kPi);
    valid &= u[0] >= kRhoFloor && thermal >= 0 && isfinite(thermal);
    if (!valid)
    {
        const double jq = src[v.src.at(8, q[0], q[1], q[2])];
        for (int c = 0; c < 8; ++c) u[c] = src[v.src.at(c, q[0], q[1], q[2])]/jq;
    }
    double out[8];
    conservedToPrimitive<double>(out, u, 1.0, v.gamma);
    for (int c = 0; c < 8; ++c) dst[v.dst.at(c, p[0], p[1], p[2])] = out[c];
}

struct AvgDownView
{
    Patch src, dst;
    int ncomp;
    int ratio[3];
};

// Restriction of fine averages to their parents (Stencil::AvgDown, applied by
// restrictTo, src/MHD_AMR.H:596-606): the mean of the r0 r1 r2 children.
__global__ void __launch_bounds__(kAmrBlock) k_amr_avgdown(AvgDownView v, const double* __restrict__ src,
                                                           double* __restrict__ dst)
{
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= v.dst.count()) return;
    int q[3];
    q[0] = v.dst.lo[0] + (int)(t % v.dst.n(0));
    q[1] = v.dst.lo[1] + (int)((t / v.dst.n(0)) % v.dst.n(1));
    q[2] = v.dst.lo[2] + (int)(t / (v.dst.n(0)*v.dst.n(1)));
    const double inv = 1.0/(double(v.ratio[0])*v.ratio[1]*v.ratio[2]);
    for (int c = 0; c < v.ncomp; ++c)
    {
        double sum = 0.0;
        for (int c2 = 0; c2 < v.ratio[2]; ++c2)
            for (int c1 = 0; c1 < v.ratio[1]; ++c1)
                for (int c0 = 0; c0 < v.ratio[0]; ++c0)
                    sum += inv*src[v.src.at(c, q[0]*v.ratio[0]+c0, q[1]*v.ratio[1]+c1, q[2]*v.ratio[2]+c2)];
        dst[v.dst.at(c, q[0], q[1], q[2])] = sum;
    }
}

struct Prolong4View
{
    Patch fine, coarse, parents;   // parents: the coarse cells whose children fill `fine`
    int ratio[3];
    double gamma;
};

// Admissible parent state: density >= RHO00 (64 ulp slack) and non-negative
// thermal energy within 64 ulp of the energy scale, amrAdmissibleParent
// (src/MHD_AMR.H:58-73).
__device__ inline bool admissibleParent(const double u[8])
{
    const double eps = 2.220446049250313e-16;
    if (!(u[0] >= kRhoFloor*(1 - 64*eps)) || !isfinite(u[0])) return false;
    double thermal = u[4], scale = fabs(u[4]);
    for (int d = 0; d < 3; ++d)
    {
        const double energy = 0.5*u[1+d]*u[1+d]/u[0] + u[5+d]*u[5+d]/(8*kPi);
        thermal -= energy; scale += energy;
    }
    return isfinite(thermal) && thermal >= -64*eps*scale;
}

// Conservative fourth-order refinement of one parent (prolongFourthOrder,
// src/MHD_AMRFourthOrder.H:253-307): the children's physical values are the
// interpolated J*U over the interpolated J; their J-weighted mean is replaced
// by the parent's J*U (times the child count over the children's J sum), and
// the deviations are scaled by the largest theta in [0, 1] (48 bisection
// steps) that keeps every child's density >= RHO00 and thermal energy >= 0.
// status[parent] = 0 when done, 1 when a child's interpolated values are not
// usable or the parent itself is not admissible.
__global__ void __launch_bounds__(kAmrBlock) k_amr_prolong4(Prolong4View v, const double* __restrict__ interp,
                                                            const double* __restrict__ fineJ,
                                                            const double* __restrict__ coarse,
                                                            double* __restrict__ out, int* __restrict__ status)
{
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= v.parents.count()) return;
    int q[3];
    q[0] = v.parents.lo[0] + (int)(t % v.parents.n(0));
    q[1] = v.parents.lo[1] + (int)((t / v.parents.n(0)) % v.parents.n(1));
    q[2] = v.parents.lo[2] + (int)(t / (v.parents.n(0)*v.parents.n(1)));
    const int r0 = v.ratio[0], r1 = v.ratio[1], r2 = v.ratio[2];
    auto child = [&](int n, int p[3]) {
        p[0] = q[0]*r0 + n % r0; p[1] = q[1]*r1 + (n / r0) % r1; p[2] = q[2]*r2 + n / (r0*r1); };
    const int nch = r0*r1*r2;
    double sumJ = 0.0, mean[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    bool available = true;
    for (int n = 0; n < nch && available; ++n)
    {
        int p[3];
        child(n, p);
        const double iJ = interp[v.fine.at(8, p[0], p[1], p[2])];
        if (!isfinite(iJ) || !(iJ > 0)) { available = false; break; }
        const double j = fineJ[v.fine.at(0, p[0], p[1], p[2])];
        sumJ += j;
        for (int c = 0; c < 8; ++c)
        {
            const double physical = interp[v.fine.at(c, p[0], p[1], p[2])]/iJ;
            if (!isfinite(physical)) available = false;
            out[v.fine.at(c, p[0], p[1], p[2])] = physical;
            mean[c] += j*physical;
        }
    }
    if (!available) { status[t] = 1; return; }
    const double cells = double(r0)*r1*r2;
    double center[8];
    for (int c = 0; c < 8; ++c)
    {
        mean[c] /= sumJ;
        center[c] = cells*coarse[v.coarse.at(c, q[0], q[1], q[2])]/sumJ;
    }
    auto admissible = [&](double theta) {
        for (int n = 0; n < nch; ++n)
        {
            int p[3];
            child(n, p);
            double u[8];
            for (int c = 0; c < 8; ++c) u[c] = center[c] + theta*(out[v.fine.at(c, p[0], p[1], p[2])] - mean[c]);
            if (theta == 0)
            {
                if (!admissibleParent(u)) return false;
                continue;
            }
            if (!(u[0] >= kRhoFloor)) return false;
            double thermal = u[4];
            for (int d = 0; d < 3; ++d) thermal -= 0.5*u[1+d]*u[1+d]/u[0] + u[5+d]*u[5+d]/(8*kPi);
            if (!isfinite(thermal) || thermal < 0) return false;
        }
        return true; };
    double theta = 1;
    if (!admissible(theta))
    {
        if (!admissible(0)) { status[t] = 1; return; }
        double lo = 0, hi = 1;
        for (int n = 0; n < 48; ++n)
        {
            const double trial = 0.5*(lo + hi);
            if (admissible(trial)) lo = trial; else hi = trial;
        }
        theta = lo;
    }
    for (int n = 0; n < nch; ++n)
    {
        int p[3];
        child(n, p);
        const double j = fineJ[v.fine.at(0, p[0], p[1], p[2])];
        for (int c = 0; c < 8; ++c)
        {
            const long long o = v.fine.at(c, p[0], p[1], p[2]);
            out[o] = j*(center[c] + theta*(out[o] - mean[c]));
        }
    }
    status[t] = 0;
}

// Sum of n per-parent flags into *count (one block).
__global__ void __launch_bounds__(kAmrBlock) k_count_flags(const int* __restrict__ flags, long long n,
                                                           int* __restrict__ count)
{
    __shared__ int part[kAmrBlock];
    int sum = 0;
    for (long long x = threadIdx.x; x < n; x += blockDim.x) sum += flags[x];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int w = kAmrBlock/2; w > 0; w >>= 1)
    {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *count = part[0];
}

// Average over the child interval [x - h/2, x + h/2] (coarse cell units,
// parent centred at 0) of the quartic whose averages over the coarse cells
// [k - 1/2, k + 1/2], k = -2..2, are the data: solve sum_k w_k <xi^n>_k =
// <xi^n>_child for n = 0..4 (Gaussian elimination with partial pivoting).
void childWeights(int ratio, int child, double out[5])
{
    if (ratio == 1) { for (int k = 0; k < 5; ++k) out[k] = k == 2 ? 1.0 : 0.0; return; }
    const double h = 1.0/ratio, x = (child + 0.5)/ratio - 0.5;
    auto avgPow = [](double a, double b, int n) {
        return (std::pow(b, n+1) - std::pow(a, n+1))/((n+1)*(b - a)); };
    double M[5][6];
    for (int n = 0; n < 5; ++n)
    {
        for (int k = 0; k < 5; ++k) M[n][k] = avgPow(k - 2 - 0.5, k - 2 + 0.5, n);
        M[n][5] = avgPow(x - h/2, x + h/2, n);
    }
    for (int col = 0; col < 5; ++col)
    {
        int piv = col;
        for (int r = col+1; r < 5; ++r) if (std::fabs(M[r][col]) > std::fabs(M[piv][col])) piv = r;
        for (int c = 0; c < 6; ++c) std::swap(M[col][c], M[piv][c]);
        for (int r = 0; r < 5; ++r)
        {
            if (r == col) continue;
            const double f = M[r][col]/M[col][col];
            for (int c = col; c < 6; ++c) M[r][c] -= f*M[col][c];
        }
    }
    for (int k = 0; k < 5; ++k) out[k] = M[k][5]/M[k][k];
}

Patch toPatch(const hc_box& b)
{
    Patch p;
    for (int d = 0; d < 3; ++d) { p.lo[d] = b.lo[d]; p.hi[d] = b.hi[d]; }
    return p;
}

bool validBox(const hc_box* b)
{
    return b && b->hi[0] >= b->lo[0] && b->hi[1] >= b->lo[1] && b->hi[2] >= b->lo[2];
}
} // namespace

void amrChildWeights(int ratio, int child, double w[5]) { childWeights(ratio, child, w); }
} // namespace hc

namespace
{
// Argument checks and the launch view shared by both refinement entry points:
// an error text, or nullptr with v filled.
const char* prolongSetup(const double* interp, const double* fine_j, const hc_box* fine_box, const double* coarse,
                         const hc_box* coarse_box, const int* ratio, double gamma, const double* fine_ju,
                         hc::Prolong4View& v)
{
    using namespace hc;
    if (!interp || !fine_j || !coarse || !fine_ju || !ratio || !validBox(fine_box) || !validBox(coarse_box))
        return "null argument or empty box";
    v = Prolong4View{};
    v.fine = toPatch(*fine_box);
    v.coarse = toPatch(*coarse_box);
    v.gamma = gamma;
    for (int d = 0; d < 3; ++d)
    {
        if (ratio[d] < 1) return "ratio must be >= 1 per direction";
        v.ratio[d] = ratio[d];
        // the fine box must be whole parents, all inside the coarse box
        if (floorDiv(fine_box->lo[d], ratio[d])*ratio[d] != fine_box->lo[d] ||
            floorDiv(fine_box->hi[d] + 1, ratio[d])*ratio[d] != fine_box->hi[d] + 1)
            return "the fine box must cover whole coarse cells";
        v.parents.lo[d] = floorDiv(fine_box->lo[d], ratio[d]);
        v.parents.hi[d] = floorDiv(fine_box->hi[d] + 1, ratio[d]) - 1;
        if (v.parents.lo[d] < coarse_box->lo[d] || v.parents.hi[d] > coarse_box->hi[d])
            return "parents outside the coarse box";
    }
    return nullptr;
}

// The device address of `scratch` on the current device, looked up once per
// device.
int* noallocScratch()
{
    constexpr int kMaxDevices = 64;
    static std::mutex guard;
    static int* address[kMaxDevices] = {};
    int dev = 0;
    HC_THROW(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(guard);
    if (!address[dev])
    {
        void* p = nullptr;
        HC_THROW(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        address[dev] = static_cast<int*>(p);
    }
    return address[dev];
}
} // namespace

extern "C" {

int hc_amr_interp4(const double* coarse, const hc_box* coarse_box, int ncomp, const int* ratio,
                   const hc_box* fine_box, double* fine, hc_stream stream)
{
    using namespace hc;
    if (!coarse || !fine || !ratio || !validBox(coarse_box) || !validBox(fine_box))
        return fail("hc_amr_interp4: null argument or empty box");
    if (ncomp < 1) return fail("hc_amr_interp4: ncomp must be >= 1");
    for (int d = 0; d < 3; ++d)
        if (ratio[d] < 1 || ratio[d] > kMaxRatio) return fail("hc_amr_interp4: ratio must be 1..8 per direction");
    try
    {
        Interp4View v{};
        v.src = toPatch(*coarse_box);
        v.dst = toPatch(*fine_box);
        v.ncomp = ncomp;
        for (int d = 0; d < 3; ++d)
        {
            v.ratio[d] = ratio[d];
            for (int c = 0; c < ratio[d]; ++c) childWeights(ratio[d], c, v.w[d][c]);
        }
        const long long n = v.dst.count();
        hipLaunchKernelGGL(k_amr_interp4, dim3((unsigned)((n + kAmrBlock - 1)/kAmrBlock)), dim3(kAmrBlock), 0,
                           static_cast<hipStream_t>(stream), v, coarse, fine);
        HC_THROW(hipGetLastError());
        return 0;
    }
    catch (const std::exception& e) { return fail(std::string("hc_amr_interp4: ") + e.what()); }
}

int hc_amr_ghosts2(const double* coarse_ju_j, const hc_box* coarse_box, const int* ratio, const hc_box* fine_box,
                   const hc_box* owned_box, double gamma, double* fine_w, hc_stream stream)
{
    using namespace hc;
    if (!coarse_ju_j || !fine_w || !ratio || !validBox(coarse_box) || !validBox(fine_box) || !validBox(owned_box))
        return fail("hc_amr_ghosts2: null argument or empty box");
    for (int d = 0; d < 3; ++d)
        if (ratio[d] < 1) return fail("hc_amr_ghosts2: ratio must be >= 1 per direction");
    try
    {
        Ghost2View v{};
        v.src = toPatch(*coarse_box);
        v.dst = toPatch(*fine_box);
        v.owned = toPatch(*owned_box);
        for (int d = 0; d < 3; ++d) v.ratio[d] = ratio[d];
        v.gamma = gamma;
        const long long n = v.dst.count();
        hipLaunchKernelGGL(k_amr_ghosts2, dim3((unsigned)((n + kAmrBlock - 1)/kAmrBlock)), dim3(kAmrBlock), 0,
                           static_cast<hipStream_t>(stream), v, coarse_ju_j, fine_w);
        HC_THROW(hipGetLastError());
        return 0;
    }
    catch (const std::exception& e) { return fail(std::string("hc_amr_ghosts2: ") + e.what()); }
}

int hc_amr_average_down(const double* fine, const hc_box* fine_box, int ncomp, const int* ratio,
                        const hc_box* coarse_box, double* coarse, hc_stream stream)
{
    using namespace hc;
    if (!fine || !coarse || !ratio || !validBox(fine_box) || !validBox(coarse_box))
        return fail("hc_amr_average_down: null argument or empty box");
    if (ncomp < 1) return fail("hc_amr_average_down: ncomp must be >= 1");
    for (int d = 0; d < 3; ++d)
    {
        if (ratio[d] < 1) return fail("hc_amr_average_down: ratio must be >= 1 per direction");
        // every child of every coarse cell must lie in the fine patch
        if ((long long)coarse_box->lo[d]*ratio[d] < fine_box->lo[d] ||
            (long long)coarse_box->hi[d]*ratio[d] + ratio[d] - 1 > fine_box->hi[d])
            return fail("hc_amr_average_down: coarse box not covered by the fine patch");
    }
    try
    {
        AvgDownView v{};
        v.src = toPatch(*fine_box);
        v.dst = toPatch(*coarse_box);
        v.ncomp = ncomp;
        for (int d = 0; d < 3; ++d) v.ratio[d] = ratio[d];
        const long long n = v.dst.count();
        hipLaunchKernelGGL(k_amr_avgdown, dim3((unsigned)((n + kAmrBlock - 1)/kAmrBlock)), dim3(kAmrBlock), 0,
                           static_cast<hipStream_t>(stream), v, fine, coarse);
        HC_THROW(hipGetLastError());
        return 0;
    }
    catch (const std::exception& e) { return fail(std::string("hc_amr_average_down: ") + e.what()); }
}


int hc_amr_prolong4(const double* interp, const double* fine_j, const hc_box* fine_box, const double* coarse,
                    const hc_box* coarse_box, const int* ratio, double gamma, double* fine_ju, int* invalid,
                    hc_stream stream)
{
    using namespace hc;
    try
    {
        Prolong4View v;
        if (const char* e = prolongSetup(interp, fine_j, fine_box, coarse, coarse_box, ratio, gamma, fine_ju, v))
            return fail(std::string("hc_amr_prolong4: ") + e);
        const long long n = v.parents.count();
        int* status = nullptr;
        HC_THROW(hipMalloc(&status, n*sizeof(int)));
        hipStream_t s = static_cast<hipStream_t>(stream);
        hipLaunchKernelGGL(k_amr_prolong4, dim3((unsigned)((n + kAmrBlock - 1)/kAmrBlock)), dim3(kAmrBlock), 0, s,
                           v, interp, fine_j, coarse, fine_ju, status);
        std::vector<int> host(n);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(host.data(), status, n*sizeof(int), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        (void)hipFree(status);
        HC_THROW(e);
        int bad = 0;
        for (long long x = 0; x < n; ++x) bad += host[x];
        if (invalid) *invalid = bad;
        if (bad) return fail("hc_amr_prolong4: invalid or unavailable parent data for conservative refinement");
        return 0;
    }
    catch (const std::exception& e) { return fail(std::string("hc_amr_prolong4: ") + e.what()); }
}

int hc_amr_prolong4_noalloc(const double* interp, const double* fine_j, const hc_box* fine_box, const double* coarse,
                            const hc_box* coarse_box, const int* ratio, double gamma, double* fine_ju, int* invalid,
                            hc_stream stream)
{
    using namespace hc;
    try
    {
        Prolong4View v;
        if (const char* e = prolongSetup(interp, fine_j, fine_box, coarse, coarse_box, ratio, gamma, fine_ju, v))
            return fail(std::string("hc_amr_prolong4_noalloc: ") + e);
        const long long n = v.parents.count();
        if (n > kNoAllocParents)
            return fail("hc_amr_prolong4_noalloc: " + std::to_string(n) + " parent cells exceed the " +
                        std::to_string(kNoAllocParents) + " of the static scratch (use hc_amr_prolong4)");
        int* flags = noallocScratch();
        int* count = flags + kNoAllocParents;
        hipStream_t s = static_cast<hipStream_t>(stream);
        // every flag of the n parents is written by the refinement kernel, and
        // the count by the sum kernel: nothing in the scratch needs clearing
        hipLaunchKernelGGL(k_amr_prolong4, dim3((unsigned)((n + kAmrBlock - 1)/kAmrBlock)), dim3(kAmrBlock), 0, s,
                           v, interp, fine_j, coarse, fine_ju, flags);
        hipLaunchKernelGGL(k_count_flags, dim3(1), dim3(kAmrBlock), 0, s, flags, n, count);
        int bad = 0;
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(&bad, count, sizeof(int), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        HC_THROW(e);
        if (invalid) *invalid = bad;
        if (bad) return fail("hc_amr_prolong4_noalloc: invalid or unavailable parent data for conservative refinement");
        return 0;
    }
    catch (const std::exception& e) { return fail(std::string("hc_amr_prolong4_noalloc: ") + e.what()); }
}

} // extern "C"
