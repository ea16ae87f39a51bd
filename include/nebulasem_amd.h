/*
 * nebulasem_amd.h — C-ABI of the MI355X-native NebulaSEM hot path.
 *
 * Two paths of NebulaSEM (dshawul/NebulaSEM) run here on one MI355X (gfx950):
 *   1. the explicit dGSEM residual + update of the compressible Euler solver
 *      (apps/euler/euler.cpp:179-287: divf -> rusanov -> div_flux + volume
 *      term -> addTemporal -> Solve), and
 *   2. the FV pressure Poisson solve (lap -> MeshMatrix -> SolveT Jacobi-PCG,
 *      src/field/field.h:3571-3704, src/solvers/solve.cpp:39-559).
 * Plain pointers and sizes only.  Every function returns an nsem_status; on
 * failure nsem_last_error() describes the cause.  Node arrays are indexed as
 * the reference indexes them: node = cell*NP + i*NPY*NPZ + j*NPZ + k
 * (src/field/dg.h:43-44), interior nodes first.
 */
#ifndef NEBULASEM_AMD_H
#define NEBULASEM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    NSEM_OK = 0,
    NSEM_EINVAL = 1,        /* bad argument                                  */
    NSEM_EUNSUPPORTED = 2,  /* input outside what this build implements     */
    NSEM_EIO = 3,           /* file could not be read or written             */
    NSEM_EDEVICE = 4,       /* HIP runtime error / no device                 */
    NSEM_ENONFINITE = 5,    /* a NaN or Inf appeared in the state            */
    NSEM_EBREAKDOWN = 6,    /* PCG breakdown (p.Ap <= 0)                      */
    NSEM_ENOTCONVERGED = 7  /* iteration limit reached before the tolerance  */
} nsem_status;

/* Message of the last failure on this thread (never NULL). */
const char* nsem_last_error(void);
/* Library version string. */
const char* nsem_version(void);

/* ------------------------------------------------------------------ mesh --
 * Replaces the reference's mesh tool (apps/mesh/meshApp.cpp:16-283) and
 * Mesh::LoadMesh + initGeomMeshFields + DG::init_geom
 * (src/field/field.cpp:95-384, src/field/dg.cpp:147-477).  Host only.     */
typedef struct nsem_mesh nsem_mesh;

typedef struct {
    uint32_t npx, npy, npz, np, npf; /* LGL points per direction, per element, per face */
    uint32_t nbcs;                   /* interior cells (gBCS)                 */
    uint32_t ncells;                 /* interior + ghost cells (gNCells)      */
    uint32_t nfacets;                /* gNFacets                              */
    uint32_t nvertices;
    uint64_t nbcs_field;             /* interior nodes (gBCSfield)            */
    uint64_t nall_field;             /* all nodes (gALLfield)                 */
    uint32_t nboundaries;            /* named boundary patches                */
} nsem_mesh_info;

/* Block description file -> grid file (.txt or .bin), as `mesh <spec> -o <grid>`
 * (meshApp.cpp:16-283).  Straight-edged single blocks. */
int nsem_grid_generate(const char* spec_path, const char* grid_path);
/* Load a grid file and build node geometry for polynomial orders order[3]
 * (the controls keys npx npy npz).  nonortho: 0 NONE 1 MINIMUM 2 ORTHOGONAL
 * 3 OVER_RELAXED (FV only).  As Mesh::LoadMesh(0) (field.cpp:95-167),
 * including grids refined by the reference's prepare tool: cells whose sides
 * are split 2:1 get merged side outlines and mortar flags
 * (mesh.cpp:161-446, 890-1016). */
int nsem_mesh_load(const char* grid_path, const int32_t order[3], int32_t nonortho,
                   nsem_mesh** out);
/* Same from a block description file, without writing the grid. */
int nsem_mesh_from_spec(const char* spec_path, const int32_t order[3], int32_t nonortho,
                        nsem_mesh** out);
/* Same from one block given directly: corners[8][3] in the reference's block
 * corner order, ndiv[3], per-edge spacing (0 linear 1 geometric 2 wall) and
 * ratio (NULL = linear), and npatch named patches, patch i given by
 * npoints[i] points (consecutive xyz triples in `points`). */
int nsem_mesh_box(const double corners[24], const uint32_t ndiv[3], const int32_t spacing[12],
                  const double ratio[12], int32_t npatch, const char* const* names,
                  const int32_t* npoints, const double* points, const int32_t order[3],
                  int32_t nonortho, nsem_mesh** out);
int nsem_mesh_write_grid(const nsem_mesh* m, const char* grid_path);
int nsem_mesh_get_info(const nsem_mesh* m, nsem_mesh_info* info);
/* Named geometry arrays (the reference's names: FO FN fN fC fI fD cC cV Jinv
 * allFaces faceIndices0/1 gFC gFN gCC gCV FOC FNC FMC vertices facet_off
 * facet_vtx cell_off cell_faces faceid_off faceid boff bfaces has_boundary
 * xgl0..2 wgl0..2 psi0..2 dpsi0..2).  kind: 0 = uint32, 1 = double. */
int nsem_mesh_array_size(const nsem_mesh* m, const char* name, uint64_t* count, int32_t* kind);
int nsem_mesh_get_array(const nsem_mesh* m, const char* name, void* dst, uint64_t count);
int nsem_mesh_boundary_name(const nsem_mesh* m, uint32_t i, char* buf, uint64_t len);
void nsem_mesh_destroy(nsem_mesh* m);

/* ---------------------------------------------------------------- fields --
 * Reads a field file of the reference (<name><step>.txt|.bin: internal values
 * as an explicit list or up to four generators uniform / gaussian / cosine /
 * cosine2 / linear / gaussian-outside / hydrostatic, then the boundary block),
 * as MeshField::read_ (src/field/field.h:1407-1565).  values: nall_field*ncomp
 * (every node, ghosts included, xyz per node for ncomp 3); bc_kind/bc_value
 * per mesh boundary patch in nsem_mesh_boundary_name order (kind -1 when the
 * file has no entry for the patch).  gravity is used by `hydrostatic`.     */
enum { NSEM_BC_NEUMANN = 0, NSEM_BC_SYMMETRY = 1, NSEM_BC_DIRICHLET = 2, NSEM_BC_ROBIN = 3 };
int nsem_field_read(const nsem_mesh* m, const char* path, int32_t ncomp, const double gravity[3],
                    double* values, int32_t* bc_kind, double* bc_value);
/* As nsem_field_read, and ROBIN patches too (BCondition's value / shape /
 * tvalue, field.h:156-240): bc_shape[patch], bc_tvalue[ncomp*patch]; a
 * ROBIN ghost node is shape value + (1 - shape)(owner + tvalue |x_ghost -
 * x_owner|) (applyExplicitBCs, field.h:2677-2680).  nsem_field_read fails
 * with NSEM_EUNSUPPORTED on a ROBIN patch, as before. */
int nsem_field_read_bc(const nsem_mesh* m, const char* path, int32_t ncomp, const double gravity[3],
                       double* values, int32_t* bc_kind, double* bc_value, double* bc_shape, double* bc_tvalue);
/* Write a field file as MeshField::write_ does (field.h:1588-1663, boundary
 * blocks as BCondition's operator<<, field.h:187-216): ".bin" paths in the
 * binary token stream (util.h:191-222), others as text with 12 significant
 * digits.  values: nbcs_field*ncomp interior values (xyz per node for ncomp
 * 3); one boundary block per patch whose bc_kind is not -1 (NSEM_BC_*),
 * value from bc_value[ncomp*patch], written when its magnitude exceeds 1e-7. */
int nsem_field_write(const nsem_mesh* m, const char* path, int32_t ncomp, const double* values,
                     const int32_t* bc_kind, const double* bc_value);
/* As nsem_field_write, with ROBIN patches (shape written when it exceeds
 * 1e-7, tvalue when its magnitude does, as BCondition's operator<<). */
int nsem_field_write_bc(const nsem_mesh* m, const char* path, int32_t ncomp, const double* values,
                        const int32_t* bc_kind, const double* bc_value, const double* bc_shape,
                        const double* bc_tvalue);
/* The refinement tree amrTree_<step>.bin / .txt that the reference's prepare
 * tool writes beside a refined grid (Mesh::Node, mesh.h:49-72, as a
 * std::vector, util.h:129-148): nodes[4*i..] = id (the cell of a leaf),
 * nchildren, cid (index of the first child node), level.  Read: pass
 * nodes = NULL to get the count in *n_out, then a buffer of capacity >= that
 * many nodes.  Replaces Prepare::refineMesh's reader / writer
 * (field.cpp:655-667, 930-936). */
int nsem_amr_tree_read(const char* path, uint32_t* nodes, uint64_t capacity, uint64_t* n_out);
int nsem_amr_tree_write(const char* path, const uint32_t* nodes, uint64_t n);

/* -------------------------------------------------------------- DG Euler --
 * Device-resident dGSEM compressible Euler problem on one GPU (gfx950).
 * Constants as apps/utils/properties.cpp:14-33 (P0 T0 cp cv viscosity Pr)
 * and the controls keys dt / gravity; buoyancy and diffusion as the euler
 * block (euler.cpp:23, 42-45, 104-131, 188-190).  diffusion != 0 adds the
 * artificial viscosity mu = rho viscosity (theta diffusivity mu / Pr) to the
 * split step and the method-of-lines steps; the derivative entry points
 * (tangent-linear, adjoint) then fail with NSEM_EUNSUPPORTED.              */
typedef struct {
    double dt;
    double P0, T0, cp, cv;
    double gravity[3];
    int32_t buoyancy;
    int32_t diffusion;
    double viscosity, Pr;
} nsem_euler_params;
typedef struct nsem_dg nsem_dg;

/* Upload geometry and build the face tables.  bc_kind[4*nb] gives, per mesh
 * boundary patch, the kinds of (rho, U, T, p) (NSEM_BC_*; scalars must be
 * NEUMANN or SYMMETRY), bc_value[6*nb] the values (rho, Ux, Uy, Uz, T, p).
 * Fails with NSEM_EDEVICE when no HIP device is present.  On a mesh with
 * 2:1 mortar faces every step and derivative entry point applies the
 * reference's mortar projections (scatter_non_conforming /
 * gather_non_conforming, field.h:2018-2248). */
int nsem_dg_create(const nsem_mesh* mesh, const nsem_euler_params* prm, const int32_t* bc_kind,
                   const double* bc_value, nsem_dg** out);
void nsem_dg_destroy(nsem_dg* d);
/* State as the reference stores it, interior nodes only (host arrays):
 * rho, U (xyz per node), T = theta', p = p'.  Setting it sets both the split
 * state and the method-of-lines state.  The two forms stay one state: the
 * entry point that advances (or reads) one form first rebuilds it on the
 * device from the other when the other was advanced last, so split steps
 * and SSP-RK3 steps may alternate.  Getting it returns the current state
 * (p' without cancellation when the method-of-lines form ran last). */
int nsem_dg_set_state(nsem_dg* d, const double* rho, const double* U, const double* T, const double* p);
int nsem_dg_get_state(nsem_dg* d, double* rho, double* U, double* T, double* p);
/* ---- derivatives of the FP64 method-of-lines step (no reference
 * counterpart; for gradient checks).  Conserved-variable arrays are 5*nb
 * doubles: blocks rho', m_x, m_y, m_z, Theta' of nb nodes each.  They act on
 * the method-of-lines state (set by nsem_dg_set_state, advanced by
 * nsem_dg_rk3_steps).  Element order: N <= 7 in 3D for the adjoint, N <= 5
 * for the tangent-linear kernels (their LDS staging).                     */
/* L(q) and its directional derivative dL = (dL/dq) dq at the current state. */
int nsem_dg_residual_tlm(nsem_dg* d, const double* dq, double* L, double* dL);
/* t = (dL/dq)^T w at the current state. */
int nsem_dg_residual_vjp(nsem_dg* d, const double* w, double* t);
/* Advance nsteps FP64 SSP-RK3 steps and carry the tangent dq (in place) along.
 * NSEM_ENONFINITE when the state or the tangent is not finite. */
int nsem_dg_rk3_tlm(nsem_dg* d, int32_t nsteps, double* dq);
/* Advance nsteps FP64 SSP-RK3 steps from q0 and return, in place of
 * lam = dJ/dq_N, the gradient dJ/dq0 = (dq_N/dq0)^T lam.  States are
 * checkpointed every `checkpoint` steps and recomputed segment by segment.
 * NSEM_ENONFINITE when the forward state or the gradient is not finite. */
int nsem_dg_rk3_adjoint(nsem_dg* d, int32_t nsteps, int32_t checkpoint, double* lam);
/* Same, returning the device time of the sweeps (forward with checkpoints,
 * recomputation, backward), host transfers of lam excluded. */
int nsem_dg_time_rk3_adjoint(nsem_dg* d, int32_t nsteps, int32_t checkpoint, double* lam, double* seconds);
/* Per-step diagnostics of the current state (split or method-of-lines),
 * reduced on the device over the interior nodes: out[6] = Courant number
 * max, min, avg (Mesh::calc_courant, field.cpp:440-448), total mass
 * sum rho cV, energy sum rho cV (g.x + |U|^2/2 + (p/P0)^(R/cp) theta cv) and
 * volume sum cV (euler.cpp:261-283, reduce_* field.h:953-1006). */
int nsem_dg_diagnostics(nsem_dg* d, double* out);
/* nsem_dg_diagnostics without any device memory of the handle's or any
 * allocation: the block partials and the six results live in a fixed
 * 49,200-byte scratch array of the library's own, one per device for the
 * life of the process, the same size for every problem.  No counterpart in
 * the reference; for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Calls that
 * overlap on one device race on the scratch. */
int nsem_dg_diagnostics_noalloc(nsem_dg* d, double* out);
/* Hydrostatic reference p_ref, rho_ref at the interior nodes (euler.cpp:121-123). */
int nsem_dg_get_background(const nsem_dg* d, double* p_ref, double* rho_ref);
/* Initial density from theta' and p' as euler.cpp:136-149. */
int nsem_euler_initial_density(const nsem_dg* d, const double* T, const double* p, double* rho);
/* nsteps of the reference's split forward-Euler update (euler.cpp:179-287):
 * rho, then p, then U and theta, each a Rusanov dGSEM residual with the
 * explicit ddt/Solve update (field.h:3740-3920, solve.cpp:563-570); with
 * prm.diffusion the U and theta fluxes carry -mu gradf<true>(U) and
 * -mu/Pr gradf<true>(theta) (field.h:3327-3372).  Waits for the steps and
 * returns NSEM_ENONFINITE when a NaN or Inf appeared in the state (the
 * reference runs on and writes NaN fields). */
int nsem_dg_split_steps(nsem_dg* d, int32_t nsteps);
/* nsteps of SSP-RK3 (Shu-Osher) on the method-of-lines form of the same
 * spatial operator, state = conserved perturbations (rho', rho U,
 * Theta' = rho theta - rho_ref T0) over the FP64 hydrostatic background.
 * precision 64: FP64 residual; 32: FP32 residual and stage increments with
 * FP64 base state and accumulation.  No reference counterpart (the
 * reference's explicit schemes all reduce to forward Euler). */
int nsem_dg_rk3_steps(nsem_dg* d, int32_t nsteps, int32_t precision);
/* Set the method-of-lines state directly (m: xyz per node); it becomes the
 * current state. */
int nsem_dg_set_conserved(nsem_dg* d, const double* rho_p, const double* m, const double* theta_p);
/* Conserved perturbations of the method-of-lines state (interior nodes). */
int nsem_dg_get_conserved(nsem_dg* d, double* rho_p, double* m, double* theta_p);
/* Same as the step functions, returning the device time of the steps. */
int nsem_dg_time_rk3(nsem_dg* d, int32_t nsteps, int32_t precision, double* seconds);
int nsem_dg_time_split(nsem_dg* d, int32_t nsteps, double* seconds);
int nsem_dg_synchronize(nsem_dg* d);
/* *all_affine = 1 when every element's Jacobian is constant (to 1e-10
 * relative), so the method-of-lines kernels read one metric per element
 * instead of one per node; 0 otherwise.  No reference counterpart (the
 * reference stores Jinv per node, field.cpp:171-384). */
int nsem_dg_affine(const nsem_dg* d, int32_t* all_affine);

/* ------------------------------------------------------------ FV Poisson --
 * Pressure Poisson of the FV solvers: Solve(lap(p, mu, true) == divf(U))
 * (apps/potential/potential.cpp:58-64; apps/piso/piso.cpp:210-213 with
 * mu = rmu) with the DIAG (Jacobi) preconditioned CG of SolveT
 * (src/solvers/solve.cpp:39-559).  Mesh loaded with order {0,0,0}.        */
typedef struct nsem_fv nsem_fv;
typedef struct {
    int32_t iterations;        /* CG iterations (all inner ones in mode 2)   */
    int32_t outer_iterations;  /* refinement steps (mode 2), else 1          */
    int32_t breakdown;
    int32_t reserved;
    double res0_ref;           /* initial residual as the reference logs it  */
    double res_ref;            /* final sqrt(|sum z^2| / |sum x^2|)          */
    double res_true;           /* final |b - A x| / |b| in FP64               */
    double seconds;            /* device time of the solve                    */
} nsem_fv_result;
/* Assemble lap(p, mu, true) with boundary conditions folded in
 * (field.h:3571-3615, 2541-2582).  mu: per cell including ghosts (NULL = 1);
 * bc_kind/bc_value per boundary patch (NEUMANN or DIRICHLET). */
int nsem_fv_create(const nsem_mesh* mesh, const double* mu, const int32_t* bc_kind,
                   const double* bc_value, nsem_fv** out);
/* The same solver from an operator assembled by the caller: ap per cell
 * (ghosts included, ncells), ano / ann per facet with the boundary
 * conditions already folded into ap and the source (applyImplicitBCs,
 * field.h:2541-2582), so couplings to ghost cells are ignored.  The source
 * is set with nsem_fv_set_rhs.  For the source-compatible field API's
 * Solve(lap(...) == ...) (solve.cpp:574-593). */
int nsem_fv_create_coeffs(const nsem_mesh* mesh, const double* ap, const double* ano, const double* ann,
                          nsem_fv** out);
void nsem_fv_destroy(nsem_fv* fv);
/* *structured: 0 general mesh (six stored couplings per cell), 1 structured
 * block numbering (three stored coupling arrays), 2 structured with every
 * coupling of a direction bitwise equal (uniform grid and coefficient: the
 * SpMV reads no coupling array). */
int nsem_fv_info(const nsem_fv* fv, uint64_t* ncells, int32_t* structured);
/* Source = BC part - divf(U) (field.h:3470-3477): U interior cells xyz; the
 * velocity boundary conditions per patch (NEUMANN, SYMMETRY or DIRICHLET). */
int nsem_fv_set_rhs_divergence(nsem_fv* fv, const double* U, const int32_t* bc_kind,
                               const double* bc_value, int32_t nboundaries);
int nsem_fv_set_rhs(nsem_fv* fv, const double* su);
int nsem_fv_get_rhs(nsem_fv* fv, double* su);
/* y = A x for interior cells (the reference's mul, field.h:2369-2420). */
int nsem_fv_apply(nsem_fv* fv, const double* x, double* y);
/* Solve A x = b in place (x: initial guess in, solution out).  mode 0: FP64
 * PCG stopped on the reference's criterion sqrt(|sum z^2|/|sum x^2|) <= tol
 * (solve.cpp:11-35, 441-444); mode 1: FP64 PCG to |b-Ax|/|b| <= tol; mode 2:
 * FP32 inner PCG with FP64 iterative refinement to |b-Ax|/|b| <= tol.
 * Returns NSEM_ENOTCONVERGED when max_it is reached first, NSEM_EBREAKDOWN
 * when p.Ap <= 0 stops the iteration.  mode 0 | NSEM_FV_KEEP_GOING: the
 * reference's loop, which does not test p.Ap and carries on with
 * alpha = sdiv(r.z, p.Ap) (solve.cpp:424-427; sdiv is 0 for a zero
 * divisor), as the field API's Solve of a singular (all-Neumann) pressure
 * system needs. */
#define NSEM_FV_KEEP_GOING 4
int nsem_fv_solve(nsem_fv* fv, double* x, int32_t mode, double tol, int32_t max_it,
                  nsem_fv_result* res);
/* Velocity correction after the solve, U -= gradf(p, true) on the interior
 * cells (apps/potential/potential.cpp:66-67; FV gradf field.h:3363-3371):
 * p interior cells (ghosts from the pressure boundary conditions given to
 * nsem_fv_create, as Solve's applyExplicitBCs sets them), U interior cells
 * xyz, updated in place. */
int nsem_fv_correct_velocity(nsem_fv* fv, const double* p, double* U);

/* ----------------------------------------------------------- measurement --
 * Device-to-device copy of `bytes` (a multiple of 16) through a plain
 * one-float4-per-thread kernel, `reps` timed launches after one warm-up:
 * the achievable HBM bandwidth the kernels are compared with, and the
 * known-size kernel the HBM byte counters are calibrated on.  No reference
 * counterpart. */
int nsem_stream_copy(uint64_t bytes, int32_t reps, double* seconds_per_copy);

#ifdef __cplusplus
}
#endif

#endif /* NEBULASEM_AMD_H */
