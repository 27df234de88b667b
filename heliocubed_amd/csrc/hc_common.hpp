// Error reporting shared by the translation units of the C ABI library.
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "../../include/heliocubed.h"

namespace hc
{
// Message of the calling thread's last failed C-ABI call.
std::string& lastError();

inline int fail(const std::string& msg)
{
    lastError() = msg;
    return -1;
}

// Fourth-order operator selection: returns an error text or nullptr.
inline const char* checkOperatorOptions(const hc_options& o)
{
    if (o.spatial_operator != 0 && o.spatial_operator != 1 && o.spatial_operator != 2)
        return "spatial_operator must be 1 (standard) or 2 (Berta WENOZ)";
    if (o.spatial_operator == 2)
    {
        if (o.spatial_order != 4) return "the Berta operator is fourth order (spatial_order 4)";
        if (!(o.berta_eta_d > 0.0)) return "berta_eta_d must be positive";
        if (o.berta_fallback != 0 && o.berta_fallback != 1) return "berta_fallback must be 0 (WENO3) or 1 (PLM)";
    }
    if (o.divergence_cleaning != 0 && o.divergence_cleaning != 1)
        return "divergence_cleaning must be 0 (off) or 1 (Powell source)";
    return nullptr;
}

// Fourth-order level (defined with the fourth-order kernels): stage
// preparation JU -> USph (ghost 5) and the box-table RHS over the level.
struct Level4;
Level4* level4Create(const hc_shell& shell, const hc_options& opt, const int tile[3], bool withTable = true);
void level4Destroy(Level4* L);
// Rebuilds the level's box table with tiles of tile[0] x tile[1] x tile[2]
// cells (radial, alpha, beta; clipped to the panel; results unchanged).
void level4Retile(Level4& L, const int tile[3]);
void level4TileShape(const Level4& L, int tile[3]);
void level4Prepare(Level4& L, const double* U0, const double* K, double a_dt, const double* inner, double* US,
                   hipStream_t s);
void level4Rhs(Level4& L, const double* US, double* K, double* U, double b_dt, hipStream_t s);
// cons[c] += scale * (inner - outer radial boundary face fluxes) of the last
// level4Rhs (the conservation sums of radialCons).
void level4RadialCons(Level4& L, double scale, double* cons, hipStream_t s);

// Child weights of the fourth-order AMR interpolation (defined with the AMR
// transfer kernels): the average over child `child` of `ratio` of the quartic
// matching the five coarse averages around the parent.
void amrChildWeights(int ratio, int child, double w[5]);

// Level reductions (defined with their kernels): the CFL geometry tables and
// reduction buffers; the CFL reduction returns a device pointer to 1 / max
// speed sum, the state check one to the first bad flat index (or -1).
struct ShellDims;
struct CflTable;
CflTable* cflCreate(const ShellDims& dims, double gamma);
void cflDestroy(CflTable* T);
const double* cflLaunch(CflTable& T, const double* JU, double coneDensity, hipStream_t s);
const long long* stateCheckLaunch(CflTable& T, const double* U, hipStream_t s);

// Inner radial ghosts from boundary-file data (defined with its kernel).
struct BcTable;
BcTable* bcCreate(const ShellDims& dims, int order, double gamma);
void bcDestroy(BcTable* T);
void bcLaunch(BcTable& T, const double* bc, double* inner, hipStream_t s);
void outflowLaunch(BcTable& T, double strength, double time, double* inner, hipStream_t s);
// Boundary-file maps onto the inner boundary (defined with their kernels):
// fillPoints (the strip of cell-centre values) and fill (native conserved
// averages of the inner ghosts).
void bcPoints(const ShellDims& d, const double* data0, int ntheta, int nphi, const double* theta, const double* phi,
              double* bc, hipStream_t s);
void bcAverages(const ShellDims& d, int order, double gamma, const double* data0, int ntheta, int nphi,
                const double* theta, const double* phi, double* inner, hipStream_t s);
} // namespace hc

#define HC_THROW(call)                                                              \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess)                                                       \
            throw std::runtime_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)
