/*
 * HelioCubed per-box MHD right-hand side on MI355X: C ABI.
 *
 * The reference (HelioCubed, C++17 on the LBNL Proto framework) evaluates the
 * ideal-MHD RHS of a six-panel cubed-sphere shell through the BoxOp member
 *   BoxOp_EulerCubedSphere::operator()(Rhs, fluxes, USph, dt, scale, counter,
 *                                      cartesianPrimitives, activeFaces)
 *   (src/BoxOp_EulerCubedSphere.H:412-421, src/BoxOp_EulerCubedSphereImplem.H:4-402)
 * called once per box per RK stage by StageLevel
 *   (src/BoxOp_EulerCubedSphereImplem.H:955-1012)
 * after preStageLevel has filled the ghost cells
 *   (src/BoxOp_PreStage.H:427-501, src/BoxOp_SecondOrderPreStage.H:6-77),
 * with the classical RK4 of src/MHD_LevelRK.H:22-29 / src/MHD_AMRRungeKutta.H:5-41.
 * The functions below replace those entry points with stream-ordered device
 * calls. All array arguments are device pointers to double unless stated.
 *
 * Memory layout (FP64, component-major like Proto BoxData, radial index i
 * fastest). Components: rho, v or rho*v (3, Cartesian), p or E, B (3, Cartesian).
 *   valid level   V[6][8][na][na][nr]            index ((((p*8+c)*na+k)*na+j)*nr+i)
 *   stage level   W[6][8][na+2g][na+2g][nr+2g]   cell (i,j,k) at ((((p*8+c)*NK+k+g)*NJ+j+g)*NI+i+g)
 *   inner ghosts  G[6][8][na+2g][na+2g][g]       Cartesian primitives of the cells i=-g..-1
 *   face fluxes   F_d[6][8][na+(d==2)][na+(d==1)][nr+(d==0)]; face index m is the low face of cell m
 * Errors: every function returns 0 on success and a negative code on error;
 * hc_last_error() returns the message of the calling thread's last error.
 */
#ifndef HELIOCUBED_H
#define HELIOCUBED_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Shell geometry: inner_radius/outer_radius/radial_stretch are the
 * reference's ParseInputs fields of the same names (src/Inputs_Parsing.H:63-65);
 * n_rad x n_ang x n_ang cells per panel. */
typedef struct hc_shell
{
    int n_rad;
    int n_ang;
    double inner_radius;
    double outer_radius;
    double radial_stretch;
} hc_shell;

/* Operator options: the hot-path subset of ParseInputs
 * (src/Inputs_Parsing.H:71-123, names kept: spatial_operator, Berta_eta_d,
 * Berta_fallback_type). spatial_order 2 selects the PLM operator
 * (src/BoxOp_SecondOrder.H:155-370; limiter 0 none, 1 or 2 minmod, 3 MC;
 * riemann_solver 1 Rusanov, 2 Roe 8-wave); spatial_order 4 the standard
 * fourth-order operator (src/BoxOp_EulerCubedSphereImplem.H:97-402; limiter
 * 0 none, 1 Limiters_4O with flattening, 2 Limiters_minmod of the point
 * values with metric cell widths (src/BoxOp_Limiters.H:360-446,
 * src/BoxOp_EulerCubedSphereImplem.H:220-245); riemann_solver 2) with
 * state_conversion 1 (src/MHD_StateConversion.H:153-316); spatial_operator 2
 * replaces the standard operator by the Berta WENOZ operator
 * (src/BoxOp_BertaCubedSphereImplem.H:1737-2076: Jameson detector with
 * threshold berta_eta_d, WENOZ with WENO3 or PLM fallback, Roe). A fourth-order
 * level uses ghost 5; its stage array W holds native conserved averages and
 * its inner ghosts are native conserved averages. precision 0 evaluates everything
 * in FP64; precision 1 (mixed) limits slopes (O2), evaluates the Berta
 * reconstruction weights and the Riemann dissipation (Rusanov speed, Roe
 * eigen-system) in FP32 while states, jumps,
 * metric terms, central fluxes, pressure recovery, divergence and the RK
 * accumulate stay FP64. */
typedef struct hc_options
{
    int spatial_order;
    int limiter;
    int riemann_solver;
    double entropy_fix;
    double gamma;
    int precision;
    int spatial_operator;   /* 0 or 1: standard operator, 2: Berta WENOZ (fourth order) */
    double berta_eta_d;     /* Berta Jameson detector threshold (reference default 0.02) */
    int berta_fallback;     /* Berta fallback reconstruction: 0 WENO3, 1 PLM */
    int divergence_cleaning;/* 1: Powell source -J div(B) (B/4pi, v, v.B/4pi) added to the RHS
                               (src/BoxOp_SecondOrder.H:317-369 at order 2,
                               src/BoxOp_DivBcleaning.H:31-168 at order 4, with
                               src/BoxOp_BertaCubedSphereImplem.H:1911-1929,2071-2074
                               for Berta); 0: off */
} hc_options;

/* A box of the shell: panel (block) 0..5 and inclusive cell range lo..hi in
 * the panel's (radial, alpha, beta) index space, like a Proto Box of an
 * MBDisjointBoxLayout (src/BoxOp_EulerCubedSphere.H:317-368). */
typedef struct hc_box
{
    int block;
    int lo[3];
    int hi[3];
} hc_box;

typedef struct hc_level hc_level;
typedef struct hc_boxop hc_boxop;
typedef struct hc_boxop2 hc_boxop2;
typedef void* hc_stream; /* hipStream_t; NULL is the default stream */

/* Builds the level geometry (radial lines, face frames, seam gather tables)
 * on the current device. Replaces the geometry caches of
 * BoxOp_EulerCubedSphere::levelGeometry (src/BoxOp_EulerCubedSphere.H:317-368)
 * and the seam plan of src/MHD_SecondOrderSeamPlan.H:51-87. */
int hc_level_create(const hc_shell* shell, const hc_options* options, hc_level** out);
void hc_level_destroy(hc_level* level);

/* Ghost width of the stage level: BoxOp_EulerCubedSphere::ghost()
 * (src/BoxOp_EulerCubedSphere.H:376-390). */
int hc_level_ghost(const hc_level* level);

/* Element counts (doubles) of the valid, stage and inner-ghost arrays, and of
 * face array d. */
size_t hc_level_valid_count(const hc_level* level);
size_t hc_level_stage_count(const hc_level* level);
size_t hc_level_inner_count(const hc_level* level);
size_t hc_level_face_count(const hc_level* level, int d);

/* Stage preparation: U = U0 + a_dt*K (K may be NULL), conversion of the
 * mapped conserved averages J*U to Cartesian primitives, seam ghost
 * interpolation, floors and radial boundary ghosts into W.
 * Replaces preStageLevel / prepareSecondOrderPrimitives
 * (src/BoxOp_PreStage.H:427-501, src/BoxOp_SecondOrderPreStage.H:6-77,103-218). */
int hc_prepare_stage(hc_level* level, const double* U0, const double* K, double a_dt,
                     const double* inner_ghosts, double* W, hc_stream stream);

/* Level RHS: K = d<JU>/dt of every valid cell from the stage level W, with
 * shared seam face states and seam flux reconciliation; face fluxes are
 * written when F0/F1/F2 are non-NULL (then the per-direction face kernels run;
 * otherwise one fused tile kernel computes every face once in LDS). If U is non-NULL it is incremented by
 * b_dt*K in the same pass (the RK stage accumulate).
 * Replaces StageLevel + the panel register reflux
 * (src/BoxOp_EulerCubedSphereImplem.H:955-1012, src/MHD_AMR.H:1417-1499). */
int hc_stage_rhs(hc_level* level, const double* W, double* K, double* U, double b_dt,
                 double* F0, double* F1, double* F2, hc_stream stream);

/* nsteps classical RK4 steps of the level in place on U (valid layout).
 * Replaces MBLevelRK4::advance as driven by HelioCubed::MBLevelRK::advance
 * (src/MHD_LevelRK.H:22-29; tableau src/MHD_AMRRungeKutta.H:8-12). */
int hc_rk4_advance(hc_level* level, double* U, const double* inner_ghosts, double dt,
                   int nsteps, hc_stream stream);

/* hc_rk4_advance captures one RK4 step (state copy and the four stage
 * preparations and RHS evaluations) into a hipGraph on its first call and
 * replays it; the graph is re-captured when U, inner_ghosts or dt change and
 * bypassed while per-kernel profiling is on. enable = 0 launches every kernel
 * individually. No reference counterpart (Proto launches per box). */
int hc_level_use_graph(hc_level* level, int enable);

/* Conservation sums (radialCons, src/BoxOp_EulerCubedSphereImplem.H:1010 as
 * StageLevel calls it with the stage weight times dt; Proto's sum restated):
 * with cons_sum a device array of 8 doubles, every stage RHS of a
 * fourth-order level (hc_stage_rhs, hc_rk4_advance) adds b_dt times the net
 * radial inflow sum over the inner and outer boundary faces of
 * (F0(inner) - F0(outer)) dxi1 dxi2 per component, so that over RK4 steps
 * cons_sum gathers the change of sum(<JU>) dxi^3 the radial boundaries cause
 * (all of it without the Powell source). NULL turns it off. The caller zeroes
 * the array; the level only adds to it. Fourth order only. */
int hc_level_set_cons_sum(hc_level* level, double* cons_sum);

/* Tiles of the fourth-order level's box table: the level evaluates its RHS as
 * a table of t0 x t1 x t2 boxes (radial x alpha x beta cells, clipped to the
 * panel; 32^3 by default, the box size of the reference's production
 * layouts). Every tile evaluates the same per-cell arithmetic on its own
 * ghosted copy, so results are identical for every tile shape; larger tiles
 * recompute fewer ghost cells and take more scratch. hc_level_set_tile(t) is
 * hc_level_set_tile_shape(t, t, t); hc_level_tile returns the edge of a cubic
 * tile (-1 for another shape), hc_level_tile_shape the three edges. These
 * fail on a second-order level (no box table). hc_level_autotune times one
 * stage RHS of the state U (inner ghosts as for hc_rk4_advance; U is not
 * modified) for the cubes 32, 64, 128 and the radial pencils 64x32x32,
 * 128x32x32, n_rad x 32 x 32, 128x64x64, n_rad x 64 x 64 and whole panels
 * (clipped to the panel), keeps the fastest and returns its shape in
 * tile3[3] (and its time per stage, may be NULL). Box layout is chosen by
 * the user in the reference (MBDisjointBoxLayout); these have no reference
 * counterpart. */
int hc_level_set_tile(hc_level* level, int tile);
int hc_level_set_tile_shape(hc_level* level, int t0, int t1, int t2);
int hc_level_tile(const hc_level* level);
int hc_level_tile_shape(const hc_level* level, int* tile3);
int hc_level_autotune(hc_level* level, const double* U, const double* inner_ghosts, int* tile3,
                      double* ms_per_stage, hc_stream stream);

/* CFL time step of the level state U (valid layout): 1 / max over the valid
 * cells of sum_d (sqrt(gamma p/rho + |B_d|^2/(4 pi rho)) + |v_d|) / width_d,
 * each cell converted to native primitives at second order. The caller
 * multiplies *dt by its CFL number, as exec/main.cpp:345-346,375-379 does.
 * Replaces BoxOp_EulerCubedSphere::dtCFL with dtInv and f_speedmax
 * (src/BoxOp_EulerCubedSphereImplem.H:878-954,1013-1131,
 * src/BoxOp_EulerCubedSphere.H:127-194) at spatial_order 4 and
 * dtCFLSecondOrderHost (src/BoxOp_SecondOrderCFL.H:99-210) at spatial_order 2.
 * cone_boundary_density > 0 bounds the density under the Alfven term at
 * spatial_order 2 (ignored at 4, like the reference); viscosity terms and the
 * constant-field test background are not part of this path. Synchronizes
 * the stream; dt is a host pointer. */
int hc_level_dt_cfl(hc_level* level, const double* U, double cone_boundary_density, double* dt,
                    hc_stream stream);

/* Inner radial ghosts from boundary-file data: bc[6][8][na+2g][na+2g] is one
 * radial strip of primitive values per ghosted angular cell, in the file's
 * units and order (rho [amu/cm^3], v_r, v_phi, v_theta [cm/s], p [pdyne/cm^2],
 * B_r, B_phi, B_theta [uG]), as MHDReader::BoxData_to_BC supplies it for
 * BC_conserved_averages 0. Writes the level's inner ghosts
 * inner[6][8][na+2g][na+2g][g] (native conserved averages at order 4,
 * Cartesian primitives at order 2), radially scaled from r_BC = inner_radius
 * to each ghost layer. Replaces the file branch of PreStagePatch
 * (src/BoxOp_PreStage.H:161-317) and, at order 2, the ghost conversion of
 * src/BoxOp_SecondOrderPreStage.H:194-214. Stream-ordered. */
int hc_level_inner_ghosts_from_bc(hc_level* level, const double* bc, double* inner_ghosts, hc_stream stream);

/* Boundary-file frames (the arithmetic of MHDReader::BoxData_to_BC,
 * src/MHDReader.H:221-331). frames: device [nframes][8][nphi][ntheta] of the
 * file variables (n, v_r, v_phi, v_theta, p, B_r, B_phi, B_theta) on the
 * longitude-latitude nodes; data0: device [8][nphi][ntheta], the map at
 * time [s]: frame floor(time / cadence) (the last one past the end, the only
 * one if nframes == 1) rotated in longitude by 360 time / (25.38 days)
 * degrees, in whole cells of the integer 360 / nphi degrees plus a linear
 * blend, and between two frames interpolated linearly in time; with
 * frame_rotate 1 each frame rotates from its own start (BC_frame_rotate).
 * nphi <= 360. Stream-ordered. */
int hc_bc_frame_at(const double* frames, int nframes, int ntheta, int nphi, double cadence_hours, int frame_rotate,
                   double time, double* data0, hc_stream stream);
/* The map onto the level's inner boundary (MHD_BoundaryAverages,
 * src/MHD_BoundaryAverages.H:17-266); theta (polar angle from +z, ascending)
 * and phi (azimuth, ascending over one turn) are device arrays of the nodes.
 * hc_level_bc_points (fillPoints): bilinear values at the strip cell centres,
 * bc[6][8][na+2g][na+2g] (the input of hc_level_inner_ghosts_from_bc; the
 * second-order default, BC_conserved_averages 0). hc_level_bc_averages
 * (fill, BC_conserved_averages 1): native conserved averages of the inner
 * ghosts, inner_ghosts[6][8][na+2g][na+2g][g], by an 8x8 Gauss-Legendre
 * rule per angular cell and a 4-point rule per radial ghost layer (the
 * midpoint at second order). Stream-ordered. */
int hc_level_bc_points(hc_level* level, const double* data0, int ntheta, int nphi, const double* theta,
                       const double* phi, double* bc, hc_stream stream);
int hc_level_bc_averages(hc_level* level, const double* data0, int ntheta, int nphi, const double* theta,
                         const double* phi, double* inner_ghosts, hc_stream stream);

/* Inner radial ghosts of the radial-outflow test boundary (init_condition_type
 * 1, f_radialoutflowIBCs, src/BoxOp_PreStage.H:5-48 as applied at :216-227):
 * rho = 700 m_p (r_BC/r)^2, p = 1e-7 (r_BC/r)^(2 gamma), radial speed
 * 5e7 cm/s less strength/300 of it times the pattern co-rotating with the
 * Carrington period, no field; time in seconds. Same output layout as
 * hc_level_inner_ghosts_from_bc. Stream-ordered. */
int hc_level_inner_ghosts_outflow(hc_level* level, double strength, double time, double* inner_ghosts,
                                  hc_stream stream);

/* Conserved-state check of the level state U (valid layout): *valid = 1 when
 * every value is finite and every density is > 0, else 0, with
 * *first_invalid (may be NULL) the smallest flat index of an offending value
 * in the valid layout (-1 when valid). Replaces
 * BoxOp_EulerCubedSphere::validConservedState
 * (src/BoxOp_EulerCubedSphereImplem.H:813-876), which the drivers call after
 * every step (exec/main.cpp:445, src/MHD_AMR.H:422-434). Synchronizes the
 * stream. */
int hc_level_valid_state(hc_level* level, const double* U, int* valid, long long* first_invalid,
                         hc_stream stream);

/* Fourth-order box operator (spatial_order 4; the standard operator with
 * limiter 0, 1 or 2, or the Berta WENOZ operator, spatial_operator 2; Roe 8-wave,
 * riemann_solver 2, or Rusanov, 1; optional Powell source) over a table of
 * boxes, evaluated for every box in the same launches. Replaces BoxOp_EulerCubedSphere::operator()
 * (src/BoxOp_EulerCubedSphere.H:412-421, src/BoxOp_EulerCubedSphereImplem.H:4-402)
 * as called per box by StageLevel (:955-1012).
 * Arrays concatenate the boxes in table order, each box component-major
 * with radial index fastest (Proto BoxData layout):
 *   USph  native conserved cell averages on the box grown by ghost 5
 *         (rho, rho v^(r,alpha,beta), E, B^(r,alpha,beta): contravariant
 *         components in d xhat / d xi), 8*(n0+10)*(n1+10)*(n2+10) per box;
 *   rhs   d<JU>/dt per unit xi^3 on the valid box, Cartesian vector
 *         components, 8*n0*n1*n2 per box; ACCUMULATED into (zero it first);
 *   F_d   face fluxes of the valid box extruded in d (optional, NULL skips).
 * hc_boxop_count returns the element count of USph (what 0), rhs (1) and
 * F_0..F_2 (2..4). Every box must lie inside its panel and span at least 4
 * cells in each direction (the width of the limiter's one-sided stencils);
 * every level needs n_ang >= 16 (the five angular ghost layers of a box at a
 * panel edge must stay inside the panel's horizon, where the metric tables
 * are defined); an empty table is rejected. */
int hc_boxop_create(const hc_shell* shell, const hc_options* options, const hc_box* boxes, int nbox,
                    hc_boxop** out);
/* Box table over several refinement levels of an AMR hierarchy (the batched
 * form of the AMR stage evaluation, src/MHD_AMR.H:1417-1499, which calls the
 * per-box operator level by level): shells[l] gives the cell counts of level
 * l (same radii and stretch), box_levels[n] the level of box n (lo/hi in
 * that level's index space). Boxes of uneven sizes and levels are evaluated
 * in the same launches; arrays are laid out as for hc_boxop_create. */
int hc_boxop_create_levels(const hc_shell* shells, int nlevels, const hc_options* options, const hc_box* boxes,
                           const int* box_levels, int nbox, hc_boxop** out);
void hc_boxop_destroy(hc_boxop* op);
long long hc_boxop_count(const hc_boxop* op, int what);
int hc_boxop_apply(hc_boxop* op, const double* USph, double* rhs, double* F0, double* F1, double* F2,
                   hc_stream stream);
/* hc_boxop_apply with the divergence scaled: rhs += scale * (-div F), the
 * a_scale argument of BoxOp_EulerCubedSphere::operator() (dxInv =
 * -a_scale/dx, src/BoxOp_EulerCubedSphereImplem.H:355); the Powell source is
 * not scaled, as there (:386-390). hc_boxop_apply is scale 1. */
int hc_boxop_apply_scaled(hc_boxop* op, const double* USph, double* rhs, double* F0, double* F1, double* F2,
                          double scale, hc_stream stream);

/* Second-order operator over a table of boxes: secondOrderOperator
 * (src/BoxOp_SecondOrder.H:155-370) as StageLevel calls it per box at order 2
 * with Cartesian primitives (src/BoxOp_EulerCubedSphereImplem.H:1003-1004;
 * PLM with limiter 0 none, 1 or 2 minmod, 3 MC; riemann_solver 1 Rusanov or 2
 * Roe; optional Powell source), FP64. Arrays concatenate the boxes in table
 * order, each box component-major with radial index fastest:
 *   W    Cartesian primitives (rho, v, p, B) on the box grown by 2,
 *        8*(n0+4)*(n1+4)*(n2+4) per box (the caller fills the ghost cells);
 *   rhs  the valid box, ACCUMULATED: rhs += scale * (-div F) (+ the Powell
 *        source), the divergence direction by direction;
 *   F_d  face fluxes of the valid box extruded in d (NULL: kept internally).
 * Without the shared panel states of a whole level every face, seam faces
 * included, takes the PLM states of the box's own cells and ghost cells.
 * hc_boxop2_count returns the element counts of W (0), rhs (1), F_0..F_2 (2..4). */
int hc_boxop2_create(const hc_shell* shell, const hc_options* options, const hc_box* boxes, int nbox,
                     hc_boxop2** out);
void hc_boxop2_destroy(hc_boxop2* op);
long long hc_boxop2_count(const hc_boxop2* op, int what);
int hc_boxop2_apply(hc_boxop2* op, const double* W, double* rhs, double* F0, double* F1, double* F2, double scale,
                    hc_stream stream);

/* Fixed AMR hierarchy of fourth-order levels advanced by the synchronous
 * classical RK4 (advanceRungeKuttaSynchronous, src/MHD_AMRRungeKutta.H:5-41):
 * per stage the stage state is restricted fine -> coarse (restrictTo,
 * src/MHD_AMR.H:597-607), every level is prepared (prepareFourthOrder,
 * src/MHD_AMRFourthOrder.H:5-187: refined patches take their ghost cells
 * from the quartic interpolation of the parent's Cartesian averages and
 * from the patches of their own level), every patch of every level is
 * evaluated in one box table (evaluate, src/MHD_AMR.H:1417-1499), the coarse
 * cells outside a finer level are refluxed with the mean fine face flux
 * (the flux registers of src/MHD_AMR.H:651-663, 916-921), and the fine RHS
 * is restricted onto the coarse RHS; after the stages the state is averaged
 * down.
 * shells[l]: cell counts of level l (same radii and stretch; the ratios
 * n_rad(l)/n_rad(l-1) and n_ang(l)/n_ang(l-1) are integers 1..4). Level 0 is
 * the whole shell. patches / patch_levels: the refined patches (levels >= 1),
 * lo/hi in their level's index space; the patches of a level share one box
 * size (a multiple of the ratio, >= 4 cells) and are aligned to it.
 * Restrictions (rejected with an error otherwise): every refined patch keeps
 * 6 cells from the radial boundaries and 8 from the panel edges (no seam
 * interpolation or radial boundary on refined levels), lies inside one patch
 * of its parent level (level >= 2), interpolates only from cells its parent
 * level defines (level 0, or a parent patch grown by 6), and the levels are
 * properly nested (every coarse-fine face borders a parent-level cell).
 * Arrays (device, FP64): the state JU / RHS K is the level-0 valid level
 * [6][8][na][na][nr] followed by every patch [8][n2][n1][n0] in table order
 * (hc_amr4_count what 0; what 2 = the level-0 part); the stage input USph is
 * the level-0 stage layout [6][8][na+10][na+10][nr+10] followed by every
 * patch grown by 5 (what 1); inner_ghosts as for hc_prepare_stage at level 0.
 * hc_amr4_prepare: USph of the restricted JU. hc_amr4_stage_rhs: K of one
 * stage from JU (restriction, preparation, evaluation, reflux, restriction).
 * hc_amr4_rk4_advance: nsteps RK4 steps of JU in place. */
typedef struct hc_amr4 hc_amr4;
int hc_amr4_create(const hc_shell* shells, int nlevels, const hc_options* options, const hc_box* patches,
                   const int* patch_levels, int npatch, hc_amr4** out);
void hc_amr4_destroy(hc_amr4* amr);
long long hc_amr4_count(const hc_amr4* amr, int what);
/* Tiles (no reference counterpart: the reference evaluates each patch on its
 * own ghosted box). The patches of every refined level are grouped into
 * rectangular tiles of whole patches of at most refined_max cells per
 * direction (default 128 x 64 x 64; the patch size gives one tile per
 * patch), and level 0 is evaluated in tiles of level0_tile (default
 * 128 x 64 x 64, clipped to the level);
 * each tile is prepared and evaluated as one box, which gives the per-patch
 * results cell for cell with fewer ghost cells. NULL keeps a setting.
 * Synchronizes the device. hc_amr4_count(what 3) is the number of refined
 * tiles. */
int hc_amr4_set_tiles(hc_amr4* amr, const int* level0_tile, const int* refined_max);
int hc_amr4_prepare(hc_amr4* amr, const double* JU, const double* inner_ghosts, double* USph, hc_stream stream);
int hc_amr4_stage_rhs(hc_amr4* amr, const double* JU, const double* inner_ghosts, double* K, hc_stream stream);
int hc_amr4_rk4_advance(hc_amr4* amr, double* JU, const double* inner_ghosts, double dt, int nsteps,
                        hc_stream stream);

/* Host-side seeded initial state (test problem data, not on the hot path):
 * config 0 = radial wind rho=(r0/r)^2, v_r=1, B_r=+-(r0/r)^2 by hemisphere
 * (sign of z), p=0.1 rho (1 + 0.2 exp(-|x-x_c|^2/w^2)), w = 0.1 (r1-r0), x_c at
 * radius (r0+r1)/2 in a seeded direction; config 1 = the same with a
 * monopole B_r=+(r0/r)^2; config 2 = config 1 without the pulse. Writes the valid J*U level and
 * the inner ghost primitives into HOST arrays. Replaces the role of
 * BoxOp_EulerCubedSphere::initialize/primToCons (src/BoxOp_Initialize.H:47-254,
 * src/BoxOp_Utilities.H:67-219) for these problems. */
int hc_init_state(const hc_shell* shell, const hc_options* options, int config,
                  unsigned long long seed, double* JU_host, double* inner_host);

/* Host-side native conserved cell averages for the fourth-order box
 * operator (test problem data): the states of hc_init_state (config 0-2),
 * config 3 = config 0 plus Alfvenic velocity/field perturbations (amplitude
 * 0.05, angular wavenumbers <= 8, dB = -dv sqrt(4 pi rho)) and a lognormal
 * density factor exp(0.05 g - 0.00125) per cell, and config 4 = config 3 plus
 * a low-beta cone (p = 1e-4 |B|^2/2 within 15 degrees of a seeded axis),
 * config 5 = the radial wind without pulse with B_r signed by four azimuthal
 * sectors and a magnetized blob (radius 0.05 (r1-r0) at a seeded point of the
 * middle radius; density x4, v_r = 2, toroidal B of amplitude 3 |B_r|),
 * averaged over each ghosted
 * cell by a quad^3 Gauss-Legendre rule in xi. Writes the hc_boxop USph layout
 * into a HOST array. */
int hc_init_native_boxes(const hc_shell* shell, const hc_options* options, int config, unsigned long long seed,
                         const hc_box* boxes, int nbox, int quad, double* USph_host);

/* Host-side initial data of a fourth-order level (HOST arrays): J-weighted
 * cell averages <J U> of the valid level (Cartesian vector components,
 * quad^3 Gauss-Legendre in xi with the point Jacobian r^2 dr/dxi det A) and
 * the native conserved averages of the inner radial ghosts
 * [6][8][na+10][na+10][5], for the states of hc_init_native_boxes. */
int hc_init_level4(const hc_shell* shell, const hc_options* options, int config, unsigned long long seed,
                   int quad, double* JU_host, double* inner_host);

/* Host-side <J U> initial data of boxes of a fourth-order level (HOST array,
 * test problem data): the J-weighted averages of hc_init_level4 on each box,
 * each box [8][n2][n1][n0] (radial index fastest), concatenated in table
 * order. hc_init_level4's valid level is this over the six whole panels. */
int hc_init_ju_boxes(const hc_shell* shell, const hc_options* options, int config, unsigned long long seed,
                     const hc_box* boxes, int nbox, int quad, double* JU_host);

/* Host-side geometry probes (no device use): cell-averaged det A of angular
 * cell (j,k), and the face-averaged cofactor row of face (d; j,k) of panel p. */
double hc_geom_cell_detA(int n_ang, int j, int k);
void hc_geom_face_normal(int n_ang, int p, int d, int j, int k, double* normal3);

/* Host-side seam gather probe: donors of ghost column (p,j,k) of the stage
 * level; returns the donor count (<= 16). */
int hc_geom_ghost_donors(int n_ang, int p, int j, int k, int* q, int* jj, int* kk, double* w);

/* Per-kernel timing with HIP events recorded on the launch stream around
 * every kernel launch of this level (enable != 0 turns it on and clears past
 * records). hc_level_profile_read returns 0 and the kernel name, the summed
 * milliseconds and the launch count of slot `slot`, or 1 past the last slot.
 * No reference counterpart: the reference times with Proto PR_TIMERS
 * (src/BoxOp_EulerCubedSphereImplem.H:20). */
int hc_level_profile(hc_level* level, int enable);
int hc_level_profile_read(hc_level* level, int slot, char* name, int name_len,
                          double* total_ms, long long* launches);

const char* hc_last_error(void);

/* AMR coarse-to-fine transfers of one patch. Patches use the Proto BoxData
 * layout: component-major, index 0 fastest, over the inclusive lo..hi of an
 * hc_box (block ignored) in the level's index space; ghost indices may be
 * negative. ratio[d] is the refinement ratio per direction (1 = unrefined).
 *
 * hc_amr_interp4: fine cell averages fine[ncomp][fine_box] from coarse
 * averages coarse[ncomp][coarse_box] by the separable fourth-order rule of
 * amrInterpolateFourthOrder (src/MHD_AMRInterpolation.H:39-161): per refined
 * direction the average over the child of the quartic matching the five
 * coarse averages around the parent (ratio <= 8). A fine cell whose donors
 * leave coarse_box is NaN, as in the reference's pass form. */
int hc_amr_interp4(const double* coarse, const hc_box* coarse_box, int ncomp, const int* ratio,
                   const hc_box* fine_box, double* fine, hc_stream stream);

/* hc_amr_ghosts2: second-order fine ghost cells of a patch
 * (src/MHD_AMR.H:1282-1393): coarse_ju_j[9][coarse_box] holds J*U (8 Cartesian
 * conserved components) and J of the coarse cells; every cell of fine_box
 * outside owned_box gets the Cartesian primitives of the tensor-quadratic
 * average of J*U divided by that of J, or of the parent cell's J*U/J when the
 * average is not a positive state (density < RHO00 or negative thermal
 * energy). Cells whose parent J is not finite, or whose stencil leaves
 * coarse_box, are left unchanged. The constant-field test background is not
 * part of this path. */
int hc_amr_ghosts2(const double* coarse_ju_j, const hc_box* coarse_box, const int* ratio, const hc_box* fine_box,
                   const hc_box* owned_box, double gamma, double* fine_w, hc_stream stream);

/* hc_amr_average_down: coarse[ncomp][coarse_box] = mean of the ratio[0] x
 * ratio[1] x ratio[2] children in fine[ncomp][fine_box] (the restriction
 * Stencil::AvgDown of restrictTo, src/MHD_AMR.H:596-606); every child must lie
 * in fine_box. */
int hc_amr_average_down(const double* fine, const hc_box* fine_box, int ncomp, const int* ratio,
                        const hc_box* coarse_box, double* coarse, hc_stream stream);

/* hc_amr_prolong4: conservative fourth-order refinement of new fine cells
 * (prolongFourthOrder, src/MHD_AMRFourthOrder.H:226-314). interp[9][fine_box]
 * is hc_amr_interp4 of the coarse J*U and J (9 components) onto fine_box,
 * fine_j[fine_box] the fine Jacobian, coarse[9][coarse_box] the coarse J*U
 * and J. fine_box must consist of whole coarse cells inside coarse_box. Each
 * parent's children get J_fine (center + theta (u - mean)): u the
 * interpolated J*U / J, mean their J-weighted mean, center the parent's J*U
 * times the child count over the children's J sum (so the children's J*U sum
 * to the count times the parent's), theta the largest value in [0, 1] (48
 * bisection steps) keeping every child's density >= RHO00 and thermal
 * energy >= 0. Synchronizes the stream; *invalid (may be NULL) counts parents
 * whose children are unusable or which are not admissible themselves, and the
 * call then fails, as the reference throws. */
int hc_amr_prolong4(const double* interp, const double* fine_j, const hc_box* fine_box, const double* coarse,
                    const hc_box* coarse_box, const int* ratio, double gamma, double* fine_ju, int* invalid,
                    hc_stream stream);

/* hc_amr_prolong4_noalloc: hc_amr_prolong4 (same arguments, results, error
 * behaviour and stream synchronisation) that allocates nothing: its
 * per-parent flags and their count live in a static device array of the
 * library (4 MiB per GPU the library runs on, held until the process ends),
 * so it calls no hipMalloc or hipFree. At most 1048576 parent cells per call
 * (more is rejected with an error; use hc_amr_prolong4). Calls that overlap
 * on one GPU race on that array. No counterpart in the reference: it is for
 * callers that cannot allocate (tight serving loops, harnesses that forbid
 * device allocation). */
int hc_amr_prolong4_noalloc(const double* interp, const double* fine_j, const hc_box* fine_box, const double* coarse,
                            const hc_box* coarse_box, const int* ratio, double gamma, double* fine_ju, int* invalid,
                            hc_stream stream);

#ifdef __cplusplus
}
#endif
#endif
