/* wbmhd: the ideal-MHD step on one AMD MI355X (gfx950), C ABI.
 *
 * Every entry point takes device pointers, 64-bit extents, doubles and a HIP
 * stream handle (hipStream_t passed as void*), enqueues its kernels on that
 * stream and returns 0, or a negative code: WB_EINVAL for arguments it rejects,
 * -(hipError_t) for a HIP launch error.  Nothing is allocated or freed: the caller
 * owns every buffer, work arrays included.  Calls are asynchronous; the caller
 * synchronises the stream before it reads a result on the host.
 *
 * Arrays are C order with the last index contiguous, one plane per variable.  A
 * haloed 2D block of nx x ny cells is X = nx+4 by Y = ny+4 (halo 2, carafe
 * mhd.py:128); a 3D block X x Y x Z likewise.
 */
#ifndef WBMHD_H
#define WBMHD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WB_EINVAL (-100000)
/* The input needs more scratch than an allocation-free entry point's static buffer. */
#define WB_ESCRATCH (-100001)

/* ---- 2D per-kernel twins: same arguments as the reference's kernels ------ */

/* Replaces kernels/mhd_primitive/mhd_primitive_pyccel.py:23-40 (mhd_primitive).
 * In: the 6 conserved planes (X, Y) and az (X+1, Y+1).  Out: prim (8, X, Y),
 * bxf (X+1, Y), byf (X, Y+1), ecc (X, Y) and *bad (device int32), the number of
 * cells with a non-positive density or pressure or a non-finite variable. */
int wb_primitive2d(const double* rho, const double* mx, const double* my, const double* mz,
                   const double* energy, const double* bz, const double* az, int64_t X, int64_t Y,
                   double dx, double dy, double b0x, double b0y, double gamma, double* prim,
                   double* bxf, double* byf, double* ecc, int32_t* bad, void* stream);

/* Replaces kernels/mhd_sweep/mhd_sweep_pyccel.py:254-261 (mhd_sweep).  limiter:
 * 0 minmod, 1 vanleer, 2 mc, 3 none (kernels/mhd_sweep/__init__.py:13).
 * bn: (X+1, Y) for axis 0, (X, Y+1) for axis 1.  flux: (7, X-3, Y) or (7, X, Y-3). */
int wb_sweep2d(const double* prim, const double* bn, int64_t X, int64_t Y, int axis, int limiter,
               double gamma, double* flux, void* stream);

/* Replaces kernels/ct_contact/ct_contact_pyccel.py:26-33 (ct_contact).  fx_*:
 * (X-3, Y), gy_*: (X, Y-3), ecc: (X, Y); ez: (X-3, Y-3) = (nx+1, ny+1). */
int wb_ct_contact(const double* fx_by, const double* gy_bx, const double* fx_rho,
                  const double* gy_rho, const double* ecc, int64_t X, int64_t Y, double* ez,
                  void* stream);

/* Replaces kernels/mhd_rate/mhd_rate_pyccel.py:30-45 (mhd_rate).  out: device
 * double[2] = largest rate, smallest density of the block's own cells. */
int wb_rate2d(const double* rho, const double* mx, const double* my, const double* mz,
              const double* energy, const double* bz, const double* az, int64_t X, int64_t Y,
              int halo, double dx, double dy, double b0x, double b0y, double gamma, double* out,
              void* stream);

/* ---- 2D stage pieces (MHD2D._fill_halos / _patch_fluxes / _advance) ------ */

/* State buffer of a 2D patch: 6 cell planes (X*Y each) then az ((X+1)*(Y+1)).
 * Replaces the halo exchange of one patch covering its level plus the domain
 * faces, carafe mhd.py:744-777.  periodic_mask / wall_mask: bit a for axis a
 * (a non-periodic axis without its wall bit is an outflow axis). */
int wb_mhd2d_fill_halos(double* state, int64_t nx, int64_t ny, int periodic_mask, int wall_mask,
                        void* stream);

/* Wall faces of a patch, mhd.py:873-886: side_mask bits 1 x-lower, 2 x-upper,
 * 4 y-lower, 8 y-upper zero flux rows 0,2,3,4,6 and the wall row of ez. */
int wb_mhd2d_walls(double* fx, double* fy, double* ez, int64_t nx, int64_t ny, int side_mask,
                   void* stream);

/* The conservative update of patchwork's FluxRegister.update as called at
 * mhd.py:911-922 (u - (dt/dx dFx + dt/dy dFy); A_z - dt E_z) from state `in` into
 * `out`; heun != 0 fuses the average of mhd.py:983-988: out = (u' + out) * 0.5.
 * dt is a device scalar (see wb_stable_dt). */
int wb_mhd2d_update(const double* in, double* out, const double* fx, const double* fy,
                    const double* ez, int64_t nx, int64_t ny, const double* dt, double dx,
                    double dy, int heun, void* stream);

/* Replaces kernels/mhd_resistive/mhd_resistive_pyccel.py:21-32 (mhd_resistive):
 * adds eta J to the B_z fluxes and ez and eta J x B to the energy fluxes, in place.
 * prim, bxf, byf as wb_primitive2d gives them; fx, fy as wb_sweep2d; ez as
 * wb_ct_contact. */
int wb_resistive2d(const double* prim, const double* bxf, const double* byf, int64_t X, int64_t Y,
                   double dx, double dy, double eta, double* fx, double* fy, double* ez,
                   void* stream);

/* Replaces kernels/mhd_conductive/mhd_conductive_pyccel.py:41-52 (mhd_conductive):
 * subtracts the heat flux along (parallel) and across (perpendicular) the field from
 * the energy fluxes fx[4], fy[4], in place. */
int wb_conductive2d(const double* prim, const double* bxf, const double* byf, int64_t X, int64_t Y,
                    double dx, double dy, double parallel, double perpendicular, double* fx,
                    double* fy, void* stream);

/* One whole forward-Euler stage of MHD2D._advance (mhd.py:889-927: primitive,
 * both sweeps, CT-contact E_z, the heat flux and the resistive terms (eta,
 * kpar / kperp, all >= 0; zeros give the ideal kernel), the wall faces, the update;
 * with heun the average of mhd.py:983-988) in a single launch from the haloed state
 * `in` (halos filled) into `out`.  FP64: the same bits as the split kernels above.
 * wall_mask bit 0: x walls, bit 1: y walls.  nxs: cells along x per workgroup;
 * tile: cells along y per workgroup (30, 62, 126 or 254; 254 runs as 126 with a
 * diffusivity; other values run as 126). */
int wb_mhd2d_stage(const double* in, double* out, int64_t nx, int64_t ny, double dx, double dy,
                   double b0x, double b0y, double gamma, double eta, double kpar, double kperp,
                   int limiter, const double* dt, int heun, int wall_mask, int32_t* bad, int nxs,
                   int tile, void* stream);

/* The same 2D stage with the reconstruction and the HLLD flux in FP32. */
int wb_mhd2d_stage_mixed(const double* in, double* out, int64_t nx, int64_t ny, double dx,
                         double dy, double b0x, double b0y, double gamma, double eta, double kpar,
                         double kperp, int limiter, const double* dt, int heun, int wall_mask,
                         int32_t* bad, int nxs, int tile, void* stream);

/* The stable step on the device, mhd.py:954,963 (3D: mhd3d.py:722-724):
 * *dt = cfl * (1 / rate[0]) from the output of wb_rate2d/wb_rate3d, shortened to
 * `remaining` when 0 < remaining < dt (mhd.py:1004-1007); if time is not NULL,
 * *time += *dt.  No host synchronisation. */
int wb_stable_dt(const double* rate, double cfl, double remaining, double* dt, double* time,
                 void* stream);

/* The stable step with the explicit diffusion limit, mhd.py:931-963 (3D:
 * mhd3d.py:717-721): *dt = cfl * min(1 / rate[0], 0.5 / (D s)) with D = max(eta,
 * cond / rate[1]), cond = (gamma - 1) * max(kappa), s = dx^-2 + dy^-2 (+ dz^-2);
 * s = 0 takes D s = rate[2] from wb_amr2d_rate (the largest over the patches).  The
 * limit is left out where D s is 0.  remaining and time as wb_stable_dt. */
int wb_stable_dt_diffusive(const double* rate, double cfl, double remaining, double eta,
                           double cond, double s, double* dt, double* time, void* stream);

/* ---- 3D stage pieces (MHD3D) ------------------------------------------- */

/* Number of doubles of a 3D state buffer: 5 cell planes, then A_x (X, Y+1, Z+1),
 * A_y (X+1, Y, Z+1), A_z (X+1, Y+1, Z). */
int64_t wb_mhd3d_state_size(int64_t nx, int64_t ny, int64_t nz);

/* Replaces MHD3D._fill_halos, mhd3d.py:448-473 (periodic or outflow axes). */
int wb_mhd3d_fill_halos(double* state, int64_t nx, int64_t ny, int64_t nz, int periodic_mask,
                        void* stream);

/* The outflow axes of MHD3D._fill_halos only (mhd3d.py:448-473): before the fused
 * stages, which wrap their reads along the periodic axes (wrap = periodic_mask). */
int wb_mhd3d_fill_outflow(double* state, int64_t nx, int64_t ny, int64_t nz,
                          int periodic_mask, void* stream);

/* Replaces MHD3D._primitive, mhd3d.py:475-509 (+ _face_fields/_cell_fields
 * :361-386): prim (8, X, Y, Z), face fields fbx (X+1,Y,Z), fby (X,Y+1,Z),
 * fbz (X,Y,Z+1); *bad counts the bad inner cells (the caller raises). */
int wb_primitive3d(const double* state, int64_t nx, int64_t ny, int64_t nz, double dx, double dy,
                   double dz, double b0x, double b0y, double b0z, double gamma, double* prim,
                   double* fbx, double* fby, double* fbz, int32_t* bad, void* stream);

/* Replaces MHD3D._sweep, mhd3d.py:511-528 (mhd_sweep in the cyclic frame):
 * flux (7, ...) in the physical layout with n-3 faces along `axis`. */
int wb_sweep3d(const double* prim, const double* bn, int64_t nx, int64_t ny, int64_t nz, int axis,
               int limiter, double gamma, double* flux, void* stream);

/* Replaces one edge direction c of MHD3D._edge_field, mhd3d.py:530-559:
 * flux_a / flux_b are the sweeps along c+1 and c+2; e gets E_c on the block's
 * c-edges in the physical layout (n_c along c, n+1 along the other two). */
int wb_edge3d(const double* prim, const double* flux_a, const double* flux_b, int64_t nx,
              int64_t ny, int64_t nz, int c, double* e, void* stream);

/* The update of mhd3d.py:653-663 (register.update x5, A_c -= dt E_c) from `in`
 * into `out`, heun as in wb_mhd2d_update (mhd3d.py:742-747); dt on the device. */
int wb_mhd3d_update(const double* in, double* out, const double* fx, const double* fy,
                    const double* fz, const double* ex, const double* ey, const double* ez,
                    int64_t nx, int64_t ny, int64_t nz, const double* dt, double dx, double dy,
                    double dz, int heun, void* stream);

/* One whole forward-Euler stage of MHD3D._advance (mhd3d.py:635-671: primitive,
 * the three sweeps, the CT-contact edge field, the update; with heun the average
 * of mhd3d.py:742-747) in a single launch, from the haloed state `in` into `out`.
 * FP64: the same bits as the split kernels above.  nxs: cells along x per
 * workgroup segment.  wrap: bit a set for a periodic axis a along which the stage
 * reads the periodic images of its halo cells itself (that halo is not read and
 * need not be filled: wb_mhd3d_fill_outflow fills the others); the halos of the
 * axes not in wrap must be filled.  *bad (device) counts bad cells. */
int wb_mhd3d_stage(const double* in, double* out, int64_t nx, int64_t ny, int64_t nz, double dx,
                   double dy, double dz, double b0x, double b0y, double b0z, double gamma,
                   int limiter, const double* dt, int heun, int32_t* bad, int nxs, int wrap,
                   void* stream);

/* The same stage with the reconstruction and the HLLD flux in FP32; the primitive
 * conversion (E -> p), the cell electric field, the CT average, the potential
 * and the update stay FP64.  waves: 2 or 3 waves per SIMD for the register
 * allocation (a launch-shape choice; the results do not depend on it). */
int wb_mhd3d_stage_mixed(const double* in, double* out, int64_t nx, int64_t ny, int64_t nz,
                         double dx, double dy, double dz, double b0x, double b0y, double b0z,
                         double gamma, int limiter, const double* dt, int heun, int32_t* bad,
                         int nxs, int waves, int wrap, void* stream);

/* Replaces the rate of MHD3D.stable_step, mhd3d.py:674-724: out = device
 * double[2], largest rate and smallest density. */
int wb_rate3d(const double* state, int64_t nx, int64_t ny, int64_t nz, double dx, double dy,
              double dz, double b0x, double b0y, double b0z, double gamma, double* out,
              void* stream);

/* The same rate for the mixed path (MHD3D.stable_step, mhd3d.py:674-724): FP64 with
 * one division per cell and the reciprocal spacings, so the rate can differ from
 * wb_rate3d's in the last bits. */
int wb_rate3d_mixed(const double* state, int64_t nx, int64_t ny, int64_t nz, double dx,
                    double dy, double dz, double b0x, double b0y, double b0z, double gamma,
                    double* out, void* stream);

/* ---- HIP graphs of whole steps -------------------------------------------------- */
/* The step of the reference (MHD2D.step mhd.py:965-994, MHD3D.step mhd3d.py:726-754)
 * enqueues only this library's kernels with dt, the time and the counters on the
 * device, so k steps are captured by bracketing their calls on a non-default stream:
 * wb_graph_capture_begin, the step's entry points, wb_graph_capture_end (which
 * instantiates the graph into *graph).  wb_graph_launch replays it `times` times on
 * `stream`.  The graph keeps the device pointers it was captured with. */
int wb_graph_capture_begin(void* stream);
int wb_graph_capture_end(void* stream, void** graph);
int wb_graph_launch(void* graph, int times, void* stream);
int wb_graph_destroy(void* graph);

/* ---- refinement estimator ---------------------------------------------------------- */

/* The second-derivative estimator of Löhner (1987) in FLASH's normalised form,
 * max over rho and p, on every interior cell of a haloed 3D state (halos filled):
 * out (nx, ny, nz); scratch: 2 * X*Y*Z doubles (rho and p of the haloed block).
 * The reference's own estimator is |J| h (mhd3d.py:771-801); this one labels the
 * learned indicator and is its fallback.  eps: FLASH's filter (0.01). */
int wb_lohner3d(const double* state, int64_t nx, int64_t ny, int64_t nz, double dx, double dy,
                double dz, double b0x, double b0y, double b0z, double gamma, double eps,
                double* scratch, double* out, void* stream);

/* The reference's estimator, current_tags: out[i, j] = 1 where the mean of |J_z| at
 * the cell's four nodes times max(dx, dy) exceeds threshold (mhd.py:1011-1041,
 * _current mhd.py:826-831), from the face fields of wb_primitive2d (halos filled);
 * out: nx*ny bytes. */
int wb_current_tags2d(const double* bxf, const double* byf, int64_t nx, int64_t ny, double dx,
                      double dy, double threshold, uint8_t* out, void* stream);

/* 3D (mhd3d.py:771-801): |J| of the cell field's curl by central differences times
 * max(dx, dy, dz) > threshold, from prim (8, X, Y, Z) of wb_primitive3d. */
int wb_current_tags3d(const double* prim, int64_t nx, int64_t ny, int64_t nz, double dx, double dy,
                      double dz, double threshold, uint8_t* out, void* stream);

/* ---- learned refinement indicator (3D CNN, BF16 MFMA, FP32 accumulation) -------- */
/* Three 3x3x3 periodic convolutions 8 -> 16 -> 16 -> 1 with ReLU, a logit per cell;
 * activations bf16 channel-last (voxel-major, z fastest), nz a multiple of 16.
 * `params` are FP32 masters, per layer W[27][16][16] (tap, ci, co; absent channels
 * zero) then bias[16] (3 x 6928 floats); `kw` the BF16 kernel weights made from
 * them by wb_cnn_prep (31104 uint16).  The reference has no learned indicator: its
 * only estimator is current_tags (mhd3d.py:771-801); these stand beside
 * wb_lohner3d, whose flags are the labels. */
int wb_cnn_prep(const float* params, uint16_t* kw, void* stream);

/* Per-channel mean and 1/std of the 8 primitives of a haloed 3D state (norm: 16
 * floats); partial: workgroups x 16 doubles. */
int wb_cnn_moments(const double* state, int64_t nx, int64_t ny, int64_t nz, double dx, double dy,
                   double dz, double b0x, double b0y, double b0z, double gamma, double* partial,
                   int workgroups, float* norm, void* stream);

/* Network input: the normalised primitives, bf16 (nx*ny*nz x 8). */
int wb_cnn_input(const double* state, int64_t nx, int64_t ny, int64_t nz, double dx, double dy,
                 double dz, double b0x, double b0y, double b0z, double gamma, const float* norm,
                 uint16_t* out, void* stream);

/* Forward: a1, a2 (n x 16 bf16) and the logits (n floats).  v_mfma_f32_16x16x32_bf16 on
 * LDS-staged input tiles; the 1-channel logit layer puts 16 voxels along y in N with
 * banded weights instead of padding its output channel. */
int wb_cnn_forward(const uint16_t* x, const float* params, const uint16_t* kw, int64_t nx,
                   int64_t ny, int64_t nz, uint16_t* a1, uint16_t* a2, float* logit,
                   int workgroups, void* stream);

/* Backward of the mean BCE-with-logits against `label` (n floats, 0/1): the
 * gradient of every parameter into grad (params' layout) and the summed loss into
 * *loss; g3: n bf16, g2, g1: n x 16 bf16 scratch; partial: workgroups x 6928 floats. */
int wb_cnn_backward(const uint16_t* x, const uint16_t* a1, const uint16_t* a2, const float* logit,
                    const float* label, const uint16_t* kw, int64_t nx, int64_t ny, int64_t nz,
                    uint16_t* g3, uint16_t* g2, uint16_t* g1, float* partial, int workgroups,
                    float* grad, float* loss, void* stream);

/* One Adam step (Kingma and Ba 2015) of n FP32 parameters; step counts from 1. */
int wb_cnn_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2,
                float eps, int step, void* stream);

/* Single layers, for tests: conv (mode 0: bf16 relu(acc + bias) into out16; 1:
 * acc + bias of output channel 0 into logit, cin 16; 2: bf16 acc * (mask > 0), the
 * backward-data pass with kernel weights K'[t][ci][co] = W[26-t][ci][co]; `workgroups`
 * is not used by conv) and the weight gradient of one 16-output layer (27 x 16 x 16 +
 * 16 floats into out, `workgroups` partials). */
int wb_cnn_conv(const uint16_t* in, int cin, const uint16_t* wk, const float* bias, int64_t nx,
                int64_t ny, int64_t nz, int mode, const uint16_t* mask, uint16_t* out16,
                float* logit, int workgroups, void* stream);
int wb_cnn_wgrad(const uint16_t* x, int cin, const uint16_t* g16, int64_t nx, int64_t ny,
                 int64_t nz, float* partial, int workgroups, float* out, void* stream);

/* ---- 2D patch hierarchy (refined, periodic) ------------------------------------ */
/* A hierarchy lives in one FP64 buffer, patch after patch (each a 2D patch state as
 * above).  `patches` is an array of {int64 offset; int32 nx, ny; double dx, dy; int32
 * walls, pad} per patch (walls: the patch's faces on domain walls, 1 x-low, 2 x-high,
 * 4 y-low, 8 y-high: the stage zeroes their fluxes but the pressure's and E_z there); the int64 tables are offsets into the buffer built on the host from the
 * geometry (wellbalanced_amd/amr_maps.py).  Replaces patchwork's exchange_halos,
 * FluxRegister.reflux and restrict as MHD2D calls them (mhd.py:744-777,889-994). */

/* One level's halo fill in one launch: ncells ghost cells (6 fields, plane sizes per
 * entry) and nnodes A_z nodes (ghosts and shared boundary nodes) copied from their
 * canonical same-level owner; the nic other ghost cells bilinear from the level below
 * (upper-cell weights iwx, iwy) and the nin other ghost nodes by the coarse node or the
 * quadratic midpoint rule (-1/8, 3/4, 3/8) per axis (kind jkx, jky; CHANGELOG.md:10-13).
 * Levels in increasing order. */
int wb_amr2d_fill_level(double* buf, const long long* cdst, const long long* csrc,
                        const long long* cpd, const long long* cps, int ncells,
                        const long long* ndst, const long long* nsrc, int nnodes,
                        const long long* idst, const long long* ipd, const long long* ic00,
                        const long long* icy, const long long* ipc, const double* iwx,
                        const double* iwy, int nic, const long long* jdst,
                        const long long* jbase, const long long* jsx, const long long* jsy,
                        const long long* jkx, const long long* jky, int nin, void* stream);

/* The outflow or wall ghosts of the patches of a level that lie on a non-periodic
 * domain face, after the level's fill (MHD2D._fill_halos, mhd.py:748-777, with the
 * rules of mhd.py:298-353): rows of 3 int32 {patch, faces (1 x-low, 2 x-high, 4 y-low,
 * 8 y-high), walls (bit 0 x walls, bit 1 y walls)}, one per such patch; axis x first,
 * then y, each over the whole haloed extent of the other axis. */
int wb_amr2d_domain_faces(double* buf, const void* patches, const int32_t* rows, int nrows,
                          void* stream);

/* One level pair's coarsening in one launch: the na covered coarse cells := mean of their
 * four fine children (arows, 5 int64 each) and the nr coarse A_z nodes := the fine node
 * on them (rrows, 2 int64 each), mhd.py:981,991. */
int wb_amr2d_coarsen(double* buf, const long long* arows, int na, const long long* rrows, int nr,
                     void* stream);

/* Regridding (MHD2D.regrid, mhd.py:1043-1056): copy into the new buffer `buf` the
 * cells (6 fields; offsets and planes in the new and the old buffer) and A_z nodes
 * that a patch of the same level held in the old buffer. */
int wb_amr2d_regrid_copy(const double* old, double* buf, const long long* cdst,
                         const long long* csrc, const long long* cpd, const long long* cps,
                         int ncells, const long long* ndst, const long long* nsrc, int nnodes,
                         void* stream);

/* Regridding: the new cells by conservative linear prolongation from their coarse
 * cell c (row length cy, plane pc; minmod slopes, sx/sy = -1 or +1 the child), and
 * the new A_z nodes bilinear from the coarse nodes at nbase (row length nsx; kinds
 * nkx/nky 1: the mean with the next node along that axis). */
int wb_amr2d_prolong(double* buf, const long long* dst, const long long* pd, const long long* c,
                     const long long* cy, const long long* pc, const long long* sx,
                     const long long* sy, int ncells, const long long* ndst,
                     const long long* nbase, const long long* nsx, const long long* nkx,
                     const long long* nky, int nnodes, void* stream);

/* Diagnostics of a hierarchy on the device (mhd.py:1059-1170).  leaf: int32 pairs
 * (patch, haloed local cell a * (ny + 4) + b) of the leaf cells.
 * wb_amr2d_leaf_values: out (9, nleaf) = density, velocity x/y/z, pressure, field
 * x/y/z and the mean current J_z of the cell's four corners (halos filled first).
 * wb_amr2d_totals: out[6] = the integrals of the 6 conserved fields over the leaf
 * cells, by a fixed two-pass tree (the same bits on every run); workspace: device
 * doubles, wb_amr2d_totals_workspace() of them.
 * wb_amr2d_divergence: out[3] = the largest |div B| times min(dx, dy) over every
 * patch cell, the largest |coarse node - fine node| over the restriction rows (pairs
 * of int64 offsets), the largest A_z difference of neighbouring nodes of the patches
 * in fine_chunks (composite divergence = out[1] / out[2]). */
int wb_amr2d_leaf_values(const double* buf, const void* patches, const int32_t* leaf, int nleaf,
                         double b0x, double b0y, double gamma, double* out, void* stream);
int wb_amr2d_totals_workspace(void);
int wb_amr2d_totals(const double* buf, const void* patches, const int32_t* leaf, int nleaf,
                    double* workspace, double* out, void* stream);
int wb_amr2d_divergence(const double* buf, const void* patches, const int32_t* chunks,
                        int nchunks, double b0x, double b0y, const long long* rows, int nrows,
                        const int32_t* fine_chunks, int nfine, double* out, void* stream);

/* The CFL rate over every patch cell (mhd.py:930-963): out = device double[2], the
 * largest rate and the smallest density; with spatch (per patch dx^-2 + dy^-2, or
 * NULL) out[2] = the largest max(eta, cond / rho) * s of a cell (its patch's
 * diffusion number, wb_stable_dt_diffusive with s = 0).  chunks: int32 pairs
 * (patch, first cell) cutting every patch's cells into runs of 256 (built on the
 * host with the other tables). */
int wb_amr2d_rate(const double* buf, const void* patches, const int32_t* chunks, int nchunks,
                  double b0x, double b0y, double gamma, double eta, double cond,
                  const double* spatch, double* out, void* stream);

/* One stage of every patch in ONE persistent launch: `workgroups` workgroups take
 * (patch, tile, segment) items (int32 triples) from *counter until `nitems` are done
 * (the counter is zeroed first); tile: 30, 62 or 126 cells along y.  FP64 parity path;
 * _mixed: FP32 flux. */
int wb_amr2d_stage(const double* in, double* out, const void* patches, const void* items,
                   int nitems, int32_t* counter, double b0x, double b0y, double gamma, double eta,
                   double kpar, double kperp, int limiter, const double* dt, int heun,
                   int32_t* bad, int nxs, int tile, int workgroups, void* stream);
int wb_amr2d_stage_mixed(const double* in, double* out, const void* patches, const void* items,
                         int nitems, int32_t* counter, double b0x, double b0y, double gamma,
                         double eta, double kpar, double kperp, int limiter, const double* dt,
                         int heun, int32_t* bad, int nxs, int tile, int workgroups, void* stream);

/* The uniform grid's fused stage with a workgroup per 16 x 16 tile and no march
 * (arguments as wb_mhd2d_stage without the launch shape; the resistive and conduction
 * terms as there): the same bits as wb_mhd2d_stage.  Replaces MHD2D._advance
 * (mhd.py:889-927). */
int wb_mhd2d_stage_tiles(const double* in, double* out, int64_t nx, int64_t ny, double dx,
                         double dy, double b0x, double b0y, double gamma, double eta,
                         double kpar, double kperp, int limiter, const double* dt, int heun,
                         int wall_mask, int32_t* bad, void* stream);
int wb_mhd2d_stage_tiles_mixed(const double* in, double* out, int64_t nx, int64_t ny, double dx,
                               double dy, double b0x, double b0y, double gamma, double eta,
                               double kpar, double kperp, int limiter, const double* dt,
                               int heun, int wall_mask, int32_t* bad, void* stream);

/* The same stage with a workgroup per 16 x 16 tile of a patch and no march: items are
 * int32 triples (patch, tile along y, tile along x), each tile inside its patch
 * (16 tile < ny, 16 x-tile < nx), one launch of nitems workgroups, the resistive and
 * conduction terms as in wb_amr2d_stage; the same bits as wb_amr2d_stage (the same
 * expressions on the same values).  Replaces, like wb_amr2d_stage, MHD2D._advance
 * on every patch of a level (mhd.py:889-927). */
int wb_amr2d_stage_tiles(const double* in, double* out, const void* patches, const void* items,
                         int nitems, double b0x, double b0y, double gamma, double eta,
                         double kpar, double kperp, int limiter, const double* dt, int heun,
                         int32_t* bad, const int32_t* rec_first, const int32_t* rec_lists,
                         double* rec, void* stream);
int wb_amr2d_stage_tiles_mixed(const double* in, double* out, const void* patches,
                               const void* items, int nitems, double b0x, double b0y,
                               double gamma, double eta, double kpar, double kperp, int limiter,
                               const double* dt, int heun, int32_t* bad,
                               const int32_t* rec_first, const int32_t* rec_lists, double* rec,
                               void* stream);
/* The fluxes the refluxing reads, recorded by the tiled stage (rec_first NULL: none):
 * rec_first[item] is the start in rec_lists of the item's list (count, then count
 * int32 pairs (kind | index << 1, slot); kind 0: x-face (x-node i, cell j) of the tile,
 * index i * 18 + j + 1, local i in [0, 16], j in [-1, 16]; kind 1: y-face (cell i, y-node
 * j), index (i + 1) * 17 + j), and rec[6 slot + f] receives the face's flux of cell field
 * f (rho, m_x, m_y, m_z, E, B_z) in FP64.  Each face is listed by one item.  Then
 * wb_amr2d_reflux_both[_mixed] is wb_amr2d_reflux's two calls (record after stage 1,
 * reflux after stage 2) in one launch after the Heun average, on the recorded values:
 * rec1 holds stage 1's fluxes, rec2 stage 2's; rows of 8 int32 (coarse patch, cell a,
 * cell b, axis, side, slot of the coarse face, slots of the two fine faces) and the same
 * groups; the dst cell gains 0.5 d1 + 0.5 d2, the same result; no face flux is
 * recomputed and no register is kept. */
int wb_amr2d_reflux_both(double* dst, const void* patches, const int32_t* rows,
                         const int32_t* groups, int ngroups, const double* rec1,
                         const double* rec2, const double* dt, void* stream);
int wb_amr2d_reflux_both_mixed(double* dst, const void* patches, const int32_t* rows,
                               const int32_t* groups, int ngroups, const double* rec1,
                               const double* rec2, const double* dt, void* stream);

/* The flux register and refluxing (FluxRegister.update with scale 0.5 dt and
 * FluxRegister.reflux, mhd.py:911-919,989-990).  For each group (first row, count) of
 * rows (13 int64: coarse patch, cell, axis, side, face, and the two fine faces) the
 * correction d = sum of dt/h (coarse flux - mean of the fine fluxes), signed by the
 * side, is recomputed from the stage input `src`.  record != 0 (after stage 1):
 * reg[6 g + f] = 0.5 d (dst untouched); record == 0 (after stage 2 and its Heun
 * average): dst cell += reg + 0.5 d.  reg: 6 doubles per group, caller-owned. */
int wb_amr2d_reflux(const double* src, double* dst, const void* patches, const long long* rows,
                    const int32_t* groups, int ngroups, double b0x, double b0y, double gamma,
                    double eta, double kpar, double kperp, int limiter, const double* dt,
                    int record, double* reg, void* stream);
int wb_amr2d_reflux_mixed(const double* src, double* dst, const void* patches,
                          const long long* rows, const int32_t* groups, int ngroups, double b0x,
                          double b0y, double gamma, double eta, double kpar, double kperp,
                          int limiter, const double* dt, int record, double* reg, void* stream);

/* The resistive terms of MHD3D (mhd3d.py:561-602) after a stage: from the stage input
 * `in` (halos filled), eta J on the block's edges (the potential of `out` loses
 * scale * dt * eta * J) and eta J x B through its faces (the energy of `out` loses
 * scale * the flux divergence), scale = 0.5 with heun (the stage output is the Heun
 * average) else 1.  work: device doubles, wb_mhd3d_resistive_workspace() of them.
 * The sum differs from the reference's single flux difference by round-off. */
int64_t wb_mhd3d_resistive_workspace(int64_t nx, int64_t ny, int64_t nz);
int wb_mhd3d_resistive(const double* in, double* out, int64_t nx, int64_t ny, int64_t nz,
                       double dx, double dy, double dz, double b0x, double b0y, double b0z,
                       double eta, const double* dt, int heun, double* work, void* stream);

/* No counterpart in the reference: wb_mhd3d_resistive for callers that cannot
 * allocate (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  The same kernels and results, with the workspace in a static device
 * array of wb_mhd3d_resistive_noalloc_limit() doubles (16 MiB per GPU, held until the
 * process ends; grids up to 64^3); a grid whose workspace is larger returns
 * WB_ESCRATCH.  The array's address on each device is looked up once, on the first
 * call with that device current (the stream must belong to the current device).  No
 * hipMalloc or hipFree.  Every call on one GPU shares the array: calls that overlap on
 * one GPU (concurrent streams) race.  Nothing in it needs to start at zero. */
int64_t wb_mhd3d_resistive_noalloc_limit(void);
int wb_mhd3d_resistive_noalloc(const double* in, double* out, int64_t nx, int64_t ny,
                               int64_t nz, double dx, double dy, double dz, double b0x,
                               double b0y, double b0z, double eta, const double* dt, int heun,
                               void* stream);

/* ---- 3D hierarchy of equal blocks (refined, periodic) ---------------------------- */
/* Every level is a set of blocks of 16^3 cells with a halo of 2, each stored as the
 * uniform 3D state of a 16^3 grid, back to back in one buffer; level L+1 refines
 * level L by 2 as an octree.  `blocks`: per block {int64 offset; int32 level, bx, by,
 * bz; double dx, dy, dz}; `lookup`: per level the block id at every block coordinate
 * (-1: none), the levels back to back; `levels`: {int32 nlev; int32 nb[6]; int32
 * base[6]} (blocks per axis and first lookup entry of each level).  Replaces
 * patchwork's exchange_halos / restrict / FluxRegister.reflux as MHD3D.step calls
 * them (mhd3d.py:448-473,635-671,726-754), all blocks of a stage in one launch. */

/* Ghost cells and edges of every block: copies from the block of the same level
 * holding them, else interpolation from the level below. */
int wb_amr3d_fill(double* buf, const void* blocks, int nblocks, const int32_t* lookup,
                  const void* levels, void* stream);

/* The same for the listed blocks only (ids: nids int32 block ids).  restricted = 1:
 * only the ghost cells the level has no block for, and the ghost edges of those cells
 * (what a stage or reflux reading its neighbours in place still reads from the ghosts;
 * list the blocks with a missing neighbour). */
int wb_amr3d_fill_list(double* buf, const void* blocks, const int32_t* ids, int nids,
                       int restricted, const int32_t* lookup, const void* levels,
                       void* stream);

/* New blocks (ids, all of one level >= 1): every cell and edge, interior included, by
 * the fill's interpolation from the level below (regrid: mhd3d.py regrid's
 * interpolation of the new fine patches). */
int wb_amr3d_prolong(double* buf, const void* blocks, const int32_t* ids, int nids,
                     const int32_t* lookup, const void* levels, void* stream);

/* The ghosts of listed regions a block's level has no block for (the only ghosts the
 * stage and the refluxing read when they read the cells around a block in place):
 * tasks = ntasks rows {int32 block, region + 32 part, parent block, 0}, region (x + 1) *
 * 9 + (y + 1) * 3 + z + 1 of the 26 around the block.  Its cells interpolated from the
 * level below, the edges of those cells outside the block's own range from their master
 * copy on the level or interpolated, as wb_amr3d_fill: part 0 all of it, part 1 the
 * cells, part 2 + c the edges of component c.  hoods: per block the 27 offsets (int64
 * doubles, -1: none) of the blocks of its level around it, index (x + 1) * 9 + (y + 1)
 * * 3 + z + 1. */
int wb_amr3d_fill_regions(double* buf, const void* blocks, const long long* hoods,
                          const int32_t* tasks, int ntasks, void* stream);

/* One level's coarsening: the covered cells of each refined block := the mean of its
 * 8 children (cells = 1 only), and its edges := the mean of the two fine edges along
 * them (every coarse copy; the restriction of the potential, mhd3d.py:740-751);
 * children: rows of 9 int32 (the block, then its 8 children, x slowest). */
int wb_amr3d_coarsen(double* buf, const void* blocks, const int32_t* children, int nrefined,
                     int cells, const int32_t* lookup, const void* levels, void* stream);

/* The CFL rate over every block cell (mhd3d.py:674-724): out = device double[2]. */
int wb_amr3d_rate(const double* buf, const void* blocks, int nblocks, double b0x, double b0y,
                  double b0z, double gamma, double* out, void* stream);
/* The same for the mixed path: one division per cell and reciprocals, as wb_rate3d_mixed. */
int wb_amr3d_rate_mixed(const double* buf, const void* blocks, int nblocks, double b0x,
                        double b0y, double b0z, double gamma, double* out, void* stream);

/* One stage of the listed blocks (ids: nids int32 block ids; null: blocks 0 ..
 * nids - 1), FP64 parity path / FP32 flux.  direct = 1: the blocks read the cells
 * around them in place from the blocks of their level holding them (lookup and levels
 * as for the fill), and from their own ghosts only where the level has no such block
 * (a restricted fill of the blocks with a missing neighbour suffices); direct = 0: from
 * their (filled) ghosts.  The results are the same.
 * waves, the mixed kernel's launch shape: 2, one 16x16 tile per block at 2 waves per
 * SIMD; 3, half a block (8x16 tiles, two workgroups per block) at 3 waves per SIMD.
 * The results do not depend on it.
 * rec_slots, rec (both or neither): per block 6 int32 slots (face 2 axis + side, side
 * 0 the low face; -1 none) and the record they index: the stage writes the fluxes it
 * solves on those faces there, 256 face cells (u, v in the two other axes, u slower) x
 * the 5 conserved fields (double) per slot, for wb_amr3d_reflux_both. */
int wb_amr3d_stage(const double* in, double* out, const void* blocks, const int32_t* ids,
                   int nids, int direct, const int32_t* lookup, const void* levels, double b0x,
                   double b0y, double b0z, double gamma, int limiter, const double* dt, int heun,
                   int32_t* bad, const int32_t* rec_slots, double* rec, void* stream);
int wb_amr3d_stage_mixed(const double* in, double* out, const void* blocks, const int32_t* ids,
                         int nids, int direct, const int32_t* lookup, const void* levels,
                         double b0x, double b0y, double b0z, double gamma, int limiter,
                         const double* dt, int heun, int32_t* bad, int waves,
                         const int32_t* rec_slots, double* rec, void* stream);

/* The flux register and refluxing, every coarse-fine block face of every level in one
 * launch after the Heun average (FluxRegister.update with scale 0.5 dt in both stages
 * and FluxRegister.reflux, mhd3d.py:650-655,745-750): per listed coarse cell, d_s = the
 * sum over its faces of dt/h (coarse flux - mean of the four fine fluxes), signed by the
 * side, from the fluxes stage s recorded (rec1: stage 1, rec2: stage 2; slots as for
 * wb_amr3d_stage: the coarse block's face toward the fine block and the fine block's
 * face; both precisions); the cell of `dst` gains 0.5 d1 + 0.5 d2 (the register of stage 1
 * added to stage 2's half).  items: {int32 fine, coarse, ci, cj, ck, dir} per
 * coarse-fine block face (dir = 2 axis + side; side 0: the coarse cells at the fine
 * block's low face, 1: at its high face); cells: {int32 n, item[3], lane[3], pad} per
 * coarse cell with such faces, its n faces as (item, lane = u * 8 + v inside the item's
 * 8 x 8 face) in increasing dir. */
int wb_amr3d_reflux_both(double* dst, const void* blocks, const void* items, const void* cells,
                         int ncells, const int32_t* rec_slots, const double* rec1,
                         const double* rec2, const double* dt, void* stream);

/* ---- 3D hierarchies of patches of any size (MHD3D on patchwork patches) -------- */
/* Each patch holds a 3D patch state as above, patch after patch in one buffer; the
 * tables are built on the host from the boxes (wellbalanced_amd/patches3d_maps.py).
 * Replaces patchwork's exchange_halos, restrict and the regrid's move as MHD3D.step
 * and MHD3D.regrid call them (mhd3d.py:448-473,726-754,803-821). */

/* dst_buf[dst[i]] = w[s] src_buf[src[s]] + w[s+1] src_buf[src[s+1]] + ... over
 * s = start[i] .. start[i] + count[i] - 1, in that order (ghost fill of a level,
 * average-down and restriction of a level pair, the copies and interpolations of a
 * regrid).  No destination may be a source of the same call. */
int wb_patches3d_gather(double* dst_buf, const double* src_buf, const long long* dst,
                        const long long* start, const int32_t* count, const long long* src,
                        const double* w, int64_t n, void* stream);

/* The new cells of a regrid: rows of 9 int64 {dst, dst plane, coarse cell, coarse x
 * and y strides, coarse plane, sx, sy, sz}: c0 + s/4 minmod slopes per axis, 5 fields
 * (the children of a coarse cell keep its integral). */
int wb_patches3d_prolong(double* buf, const long long* rows, int64_t n, void* stream);

/* The flux register and refluxing (FluxRegister.update with scale 0.5 dt and
 * FluxRegister.reflux, mhd3d.py:650-655,745-750) from the stage's sweep arrays: per
 * refluxed cell {offset, plane, first face, face count}; per face 12 int64 {axis, side,
 * coarse flux offset and row stride, four fine ones} and the coarse spacing.  record:
 * reg[5 g + f] = half the correction; else the cell of `out` gains reg plus half the
 * correction of this stage. */
int wb_patches3d_reflux(double* out, const double* sweeps, const long long* cells,
                        const long long* faces, const double* spacing, int64_t ncells,
                        const double* dt, int record, double* reg, void* stream);

/* Replaces MHD3D._resistive (mhd3d.py:561-602) inside the split stage: from the haloed
 * face fields f0 (X+1,Y,Z), f1, f2 of wb_primitive3d, eta J is added to the edge fields
 * e0..e2 of wb_edge3d and eta (J x B)_s to the energy row of the sweeps s0..s2 of
 * wb_sweep3d on the block faces, before wb_mhd3d_update. */
int wb_resistive3d(const double* f0, const double* f1, const double* f2, int64_t nx, int64_t ny,
                   int64_t nz, double dx, double dy, double dz, double eta, double* s0,
                   double* s1, double* s2, double* e0, double* e1, double* e2, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* WBMHD_H */
