/*
 * swirlc.h -- C ABI of the MI355X (gfx950) compressible Navier-Stokes RHS and
 * Runge-Kutta step that replaces Swirl-C's TensorFlow hot path.
 *
 * The reference has no C ABI: its seam is the Python closure
 *   rhs(replica_id, conservative, helper_vars) -> FlowFieldMap
 * built by equations/compressible_navier_stokes.py:46-270 and consumed by
 * numerics/time_integration/time_integration.py:35-90.  Each entry point below
 * states the reference function it stands in for.  The Python host package
 * (swirlc_amd) binds these with ctypes and keeps the reference's Python call
 * surface above them.
 *
 * Conventions
 *   - A field is one fp32 device array shaped [nz+2h][nx+2h][ny+2h] (y
 *     fastest), the reference's FlowFieldVar layout (common/types.py:31-33).
 *     Fields are passed as host arrays of device pointers, in the order
 *     rho, rho_u, rho_v, rho_w, rho_e (common/types.py:72), then transported
 *     scalars rho_<name>.
 *   - Every call is asynchronous on the given HIP stream, performs no host
 *     synchronisation and no allocation, and may be captured in a HIP graph.
 *   - Errors: a nonzero sc_status; sc_last_error() returns the message of the
 *     calling thread's last failure.  Invalid configurations are rejected by
 *     sc_ctx_create, mirroring the reference's trace-time ValueError /
 *     NotImplementedError.
 *   - A context must not be used from two host threads at once.
 */
#ifndef SWIRLC_H_
#define SWIRLC_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int sc_status;
#define SC_OK 0
#define SC_ERR_INVALID 1     /* invalid argument or configuration */
#define SC_ERR_UNSUPPORTED 2 /* valid in the reference's API, not supported */
#define SC_ERR_HIP 3         /* HIP runtime error */

#define SC_MAX_FIELDS 7      /* 5 base conserved fields + up to 2 scalars */
#define SC_MAX_HALO 4

/* equations/compressible_navier_stokes.py:39-43 (InterpolationSchemes). */
#define SC_WENO_3 0
#define SC_WENO_5 1
#define SC_MUSCL 2
/* numerics/riemann_solver/riemann_solver.py:24-27 (NumericFluxScheme). */
#define SC_HLL 0
#define SC_HLLC 1
/* boundary/bc_types.py:22-27 (BoundaryCondition).  SC_BC_NONE marks a
 * dimension with no BC (a bc dict without that dim, boundary.py:58-60): its
 * halo cells are left as they are; both faces of the dim must carry it. */
#define SC_BC_NONE (-1)
#define SC_BC_PERIODIC 0
#define SC_BC_DIRICHLET 1
#define SC_BC_NEUMANN 2
#define SC_BC_REFLECTIVE 3
/* Runge-Kutta stage update forms (numerics/time_integration/rk3.py:69-83,
 * rk2.py:67-86), with L the RHS of uk:
 *   SC_RK_EULER: out = uk + L*dt
 *   SC_RK_BLEND: out = u0*c0 + (uk + L*dt)*c1
 *   SC_RK_AVG:   out = 0.5*(u0 + (uk + dt*L))                                 */
#define SC_RK_EULER 1
#define SC_RK_BLEND 2
#define SC_RK_AVG 3

/* One face of one field's cell-average boundary condition
 * (cfg.bc['cell_averages'][var][dim][face], core/parameter.py:112-130). */
typedef struct sc_bc_face {
  int type;                 /* SC_BC_* */
  float value;              /* DIRICHLET scalar / NEUMANN increment per cell */
  int n_layer_values;       /* DIRICHLET list of per-layer scalars (0 = none) */
  float layer_values[SC_MAX_HALO];
  const float* layer_planes;/* DIRICHLET list of per-layer planes (device, or
                               NULL): [h][A][B], A,B = full extents of the two
                               other axes in [z][x][y] order                  */
} sc_bc_face;

typedef struct sc_grid {
  int nx, ny, nz;           /* interior cells (core_nx, core_ny, core_nz) */
  int halo;                 /* halo_width */
  double dx, dy, dz;
  /* Cell-centre coordinates incl. halos (host arrays of n+2h values, built as
   * common/utils.py:41-89 does); only read when the gravity vector is a unit
   * vector (gravitational potential, generic.py:282-319).  May be NULL
   * otherwise. */
  const float* x_full;
  const float* y_full;
  const float* z_full;
} sc_grid;

/* A Dirichlet value on a boundary face of the domain (the faces between halo
 * and interior cells): cfg.bc['intercell_fluxes'][type][var][dim][side]
 * (boundary/boundary.py:89-183) and cfg.bc['cell_faces'][var][dim][side]
 * (boundary/boundary.py:230-461).  A scalar `value`, or with `plane` set a
 * value per cell of the boundary plane (the reference's
 * tensor_scatter_1d_update_global of a plane): a device array [A][B] over
 * the full extents of the two other axes in [z][x][y] order (x faces
 * [nz+2h][ny+2h], y faces [nz+2h][nx+2h], z faces [nx+2h][ny+2h]), kept
 * alive by the caller while the context is used. */
#define SC_MAX_BOUNDARY_VALUES 64
#define SC_MAX_FACE_VALUES 8      /* cell-face values per (dim, side) */
#define SC_BV_FLUX_CONVECTIVE 0   /* var: conserved field index */
#define SC_BV_FLUX_DIFFUSIVE 1
#define SC_BV_FLUX_TOTAL 2
#define SC_BV_FACE_CONSERVED 3    /* var: conserved field index 1.. (not rho) */
#define SC_BV_FACE_PRIMITIVE 4    /* var: 1 u, 2 v, 3 w, 4 e, 5 p, 6 h, 7+s scalar s */
typedef struct sc_boundary_value {
  int kind;                 /* SC_BV_* */
  int dim, side;            /* 0 x, 1 y, 2 z; 0 low, 1 high */
  int var;
  float value;
  const float* plane;       /* device plane of values, or NULL */
} sc_boundary_value;

typedef struct sc_params {
  int n_fields;             /* 5 + number of transported scalars */
  int interpolation;        /* SC_WENO_3 | SC_WENO_5 | SC_MUSCL */
  int numeric_flux;         /* SC_HLL | SC_HLLC */
  int include_diffusion;
  double nu, pr;            /* transport_parameters (physics/transport/simple.py) */
  int gravity_source;       /* 'gravity' in source_functions */
  double g[3];              /* cfg.g (unit vector or zero) */
  /* bc[f][d][s]: field f, dim d (0 x, 1 y, 2 z), side s (0 low, 1 high). */
  sc_bc_face bc[SC_MAX_FIELDS][3][2];
  /* bit f set: the BC of field f is given on its primitive (e.g. 'u' for
   * rho_u), boundary.py:516-531. */
  int prim_bc_mask;
  /* Face and flux values; cell-face values of one (dim, side) are applied
   * in list order, rho never (the reference's face update then drops p and
   * h).  Cell-face values require include_diffusion == 0. */
  int n_boundary_values;
  sc_boundary_value boundary_values[SC_MAX_BOUNDARY_VALUES];
} sc_params;

typedef struct sc_ctx sc_ctx;

/* Validates cfg and builds the per-grid context (scratch RHS accumulator,
 * potential-energy tables).  Stands in for get_rhs_fn
 * (equations/compressible_navier_stokes.py:46-67) and PhysicsModels
 * (physics/physics_models.py:45-92): invalid scheme names, missing gravity
 * direction and halo widths too small for the stencil fail here, as do
 * grids with nx >= 2^20 or (nx + 2 * halo) * (ny + 2 * halo) >= 2^30
 * (SC_ERR_UNSUPPORTED: the y sweep addresses a z plane with 32-bit byte
 * offsets). */
sc_status sc_ctx_create(const sc_grid* grid, const sc_params* params,
                        sc_ctx** out);
sc_status sc_ctx_destroy(sc_ctx* ctx);

/* Fills the halo cells of every field in place from the cell-average BCs.
 * boundary.update_conservative_cell_averages (boundary/boundary.py:464-539)
 * with swirl_lm's halo fill: dims x, y, z in order, each over full planes.
 * Fields with a primitive BC get rho*prim in their halos; their interior is
 * left untouched (the RHS applies the reference's rho*(q/rho) rounding). */
sc_status sc_fill_halos(sc_ctx* ctx, float* const* q, void* stream);

/* Right-hand side of the conserved fields (the `rhs` closure,
 * equations/compressible_navier_stokes.py:69-268): fills q's halos in place,
 * then writes rhs over the interior cells (halo cells of rhs are untouched). */
sc_status sc_rhs(sc_ctx* ctx, float* const* q, float* const* rhs, void* stream);

/* One Runge-Kutta stage: fills uk's halos, evaluates L = rhs(uk) and writes
 * the stage update (SC_RK_* form above) over interior cells of out.
 * rk3.integrate / rk2.integrate stage bodies (rk3.py:87-119, rk2.py:67-86).
 * dt_dev points to one float in device memory.  out may alias u0, not uk. */
sc_status sc_rk_stage(sc_ctx* ctx, int form, float c0, float c1,
                      const float* dt_dev, const float* const* u0,
                      float* const* uk, float* const* out, void* stream);

/* No counterpart in the reference: sc_rk_stage for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  Identical results (same kernels); the
 * RHS accumulator lives in a file-scope 128 MiB device array, one per GPU,
 * held until the process ends, instead of the context's scratch fields.
 * Limit: n_fields * ((nx+2h)(ny+2h)(nz+2h) rounded up to 64) <= 33554432
 * floats (e.g. 5 fields of a 180^3 grid with halo 3), else
 * SC_ERR_UNSUPPORTED.  The
 * context's device must be current.  Allocates nothing; needs no zeroing
 * (the x sweep writes every interior accumulator cell before it is read).
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (two streams, two threads) race. */
sc_status sc_rk_stage_noalloc(sc_ctx* ctx, int form, float c0, float c1,
                              const float* dt_dev, const float* const* u0,
                              float* const* uk, float* const* out, void* stream);

/* Hydrostatic reference state (helper_var 'p_ref' and 'rho_ref',
 * equations/compressible_navier_stokes.py:273-335): full-size device fields
 * used by the following sc_rhs / sc_rk_stage / sc_sweep_only /
 * sc_hllc_branch_codes calls until cleared with two NULLs.  p_dyn = p - p_ref
 * (zero in every halo cell) is reconstructed to the faces and replaces p in
 * the momentum flux (and, for HLLC, in the wave speeds and star states,
 * hllc.py:109-120); d_rho = rho - rho_ref replaces rho in the gravity source.
 * The caller keeps the fields alive.  Any number of fields (transported
 * scalars included). */
sc_status sc_ctx_set_reference_state(sc_ctx* ctx, const float* p_ref, const float* rho_ref);

/* Maximum over interior cells of max_d(|u_d| + sqrt(gamma p / rho)) written to
 * *out_dev, and dt = cfl * min(dx,dy,dz) / max written to *dt_dev (if not
 * NULL).  No reference counterpart (the reference takes a fixed cfg.dt); the
 * wave speed uses the reference EOS (generic.py:27-57,149-154). */
sc_status sc_max_wave_speed(sc_ctx* ctx, const float* const* q, float cfl,
                            float* out_dev, float* dt_dev, void* stream);

/* Flags of sc_rk_stage_ex. */
#define SC_STAGE_HALOS_READY 1 /* uk's halo cells already hold their BC values */
#define SC_STAGE_FILL_OUT 2    /* leave out with its halo cells filled */

/* sc_rk_stage for a sequence of stages (no reference counterpart; the
 * reference's rk3.integrate refills halos at every RHS evaluation, which this
 * does too unless told otherwise).  SC_STAGE_HALOS_READY: uk's halos are not
 * filled again (the caller states they hold the fill of uk's interior, e.g.
 * uk was the out of a call with SC_STAGE_FILL_OUT).  SC_STAGE_FILL_OUT: out's
 * halo cells are filled before the call returns (on all-periodic grids whose
 * extents are all >= 2h the last sweep writes each output cell's periodic
 * images, the same values the fill copies; otherwise sc_fill_halos runs on
 * out).  speed_dev (or NULL): a float in device memory into which the wave
 * speed of sc_max_wave_speed over out's interior cells is max-accumulated
 * (compared as its bits; clear it with sc_cfl_dt or to 0 first) - the last
 * sweep computes it from the values it writes, so a time step's stage 3 leaves
 * the next step's CFL maximum without re-reading the state, equal to
 * sc_max_wave_speed on out bit for bit. */
sc_status sc_rk_stage_ex(sc_ctx* ctx, int form, float c0, float c1, const float* dt_dev,
                         const float* const* u0, float* const* uk, float* const* out,
                         int flags, float* speed_dev, void* stream);

/* dt = cfl * min(dx,dy,dz) / *speed_dev written to *dt_dev, then *speed_dev
 * cleared to 0 (for the next sc_rk_stage_ex accumulation).  No reference
 * counterpart. */
sc_status sc_cfl_dt(sc_ctx* ctx, float* speed_dev, float cfl, float* dt_dev, void* stream);

/* Face states of one field: (neg, pos) = left / right states of face i-1/2
 * stored at index i along dim (0 x, 1 y, 2 z) of the [z][x][y] device array
 * with extents[3] = {Nx, Ny, Nz} (halos included).  scheme SC_WENO_5 /
 * SC_WENO_3 is swirl_lm's interpolation.weno(v, dim, k=3 / 2), SC_MUSCL its
 * flux_limiter(v, dim, MUSCL), as called by _face_interpolation
 * (equations/compressible_navier_stokes.py:536-615); entries whose stencil
 * leaves the array keep the cell value (WENO) or use a zero slope (MUSCL).
 * v, neg and pos must be distinct. */
sc_status sc_face_interp(const float* v, const int* extents, int dim, int scheme,
                         float* neg, float* pos, void* stream);

/* Dirichlet overwrite of one boundary plane, coordinate[dim] == index, of a
 * [z][x][y] device array of extents {Nx, Ny, Nz} with value: the plane
 * write of update_fluxes / update_flux_1d (boundary/boundary.py:89-183). */
sc_status sc_fill_plane(float* f, const int* extents, int dim, int index, float value,
                        void* stream);

/* sc_fill_plane with one value per plane cell (a plane-valued update of
 * tensor_scatter_1d_update_global, called at boundary/boundary.py:144,281,
 * 302,324): `values` is a device array [A][B] over the full extents of the two
 * other axes in [z][x][y] order: dim 0 [Nz][Ny], dim 1 [Nz][Nx], dim 2
 * [Nx][Ny]. */
sc_status sc_fill_plane_values(float* f, const int* extents, int dim, int index,
                               const float* values, void* stream);

/* Whole-array steps of the face BC update _face_boundary_update_1d
 * (boundary/boundary.py:230-337) in IEEE fp32 over n elements:
 * op 0 out = divide_no_nan(a, b), op 1 out = a * b, op 2 out = (1 / b) * a
 * (conservative_to_primitive, common/utils.py:145-183). */
sc_status sc_face_bc_elementwise(int op, const float* a, const float* b, float* out,
                                 long long n, void* stream);

/* numerics/interpolation.py:21-53 linear_interpolation: out[i] =
 * 0.5 * (v[i-1] + v[i]) on face i-1/2 along dim, v[-1] taken as 0. */
sc_status sc_linear_interp(const float* v, const int* extents, int dim, float* out,
                           void* stream);

/* Numerical flux across face i-1/2 (stored at i) normal to dim from given
 * face states, with the context's flux scheme: hll.hll_convective_flux
 * (numerics/riemann_solver/hll.py:36-175) or hllc.hllc_convective_flux
 * (hllc.py:33-257), the physical flux being _compute_physical_flux_1d
 * (equations/compressible_navier_stokes.py:338-427).  cons_*: the n_fields
 * conservative face states; prim_*: rho, u, v, w, e, p, h, then the
 * scalars' primitives; pdyn_*: helper variable p_dyn on the faces or both
 * NULL.  Full-size device arrays of the context's grid; flux: n_fields
 * outputs.  The gravitational potential is the context's (cfg.g). */
sc_status sc_riemann_flux(sc_ctx* ctx, int dim, const float* const* cons_neg,
                          const float* const* cons_pos, const float* const* prim_neg,
                          const float* const* prim_pos, const float* pdyn_neg,
                          const float* pdyn_pos, float* const* flux, void* stream);

/* physics/diffusion.py:33-67 shear_stress with fluid.py:31-111 strain rate
 * and TransportSimple (simple.py:44-79, constant kinematic nu):
 * tau[3 i + j] = 2 rho_face[j] nu S_ij on the j faces (face i-1/2 stored at
 * i).  u, v, w: cell velocities; rho_face[j]: density on the j faces; h:
 * dx, dy, dz; full [z][x][y] device arrays of extents {Nx, Ny, Nz}. */
sc_status sc_shear_stress(const float* u, const float* v, const float* w,
                          const float* const* rho_face, const int* extents, const double* h,
                          double nu, float* const* tau, void* stream);

/* physics/diffusion.py:76-109 single_component_heat_flux with
 * TransportSimple (simple.py:81-108): flux[d] = -kappa_f dT/dx_d on the d
 * faces, T = p / (rho R), kappa = (1 / pr) cp nu rho linearly interpolated. */
sc_status sc_heat_flux(const float* rho, const float* p, const int* extents, const double* h,
                       double nu, double pr, float* const* flux, void* stream);

/* boundary/boundary.py:34-86 exchange_halos for one field of one replica:
 * fills the halo layers of u ([nz+2h][nx+2h][ny+2h], extents = interior
 * {nx, ny, nz}, h = halo) in place from bc[2 * dim + side] (dim 0 x, 1 y,
 * 2 z; side 0 low, 1 high), x then y then z as swirl_lm's exchange. */
sc_status sc_exchange_halos(float* u, const int* extents, int halo, const sc_bc_face* bc,
                            void* stream);

/* Debug/parity: HLLC branch code of every face normal to dim (0 x, 1 y, 2 z),
 * codes[i] for the face i-1/2 stored at cell i of the full array, interior
 * faces only (hllc.py:227-243: 0 F_L, 1 F*_L, 2 F*_R, 3 F_R).  q's halos
 * must already be filled. */
sc_status sc_hllc_branch_codes(sc_ctx* ctx, const float* const* q, int dim,
                               uint8_t* codes, void* stream);

/* Profiling: launches only the sweep kernel of direction dim (0 x, 1 y, 2 z)
 * of an sc_rk_stage call (form 0 = plain RHS into the context's scratch), on
 * uk whose halos are already filled.  A stage runs x, y, z in that order: the
 * x sweep writes the scratch RHS, the y sweep updates it, the z sweep
 * finishes it (and writes out for form != 0).
 * No reference counterpart: it exposes one kernel of the fused stage so
 * its duration can be timed in isolation. */
sc_status sc_sweep_only(sc_ctx* ctx, int dim, int form, float c0, float c1,
                        const float* dt_dev, const float* const* u0,
                        float* const* uk, float* const* out, void* stream);

/* Profiling: sc_sweep_only whose kernel launch records its own start and stop
 * times into the HIP events ev_start / ev_stop (hipExtLaunchKernel: no extra
 * packets between kernels), so a kernel's duration inside a step sequence
 * can be read with hipEventElapsedTime.  flags (SC_STAGE_FILL_OUT only: the
 * periodic images the z sweep writes on all-periodic grids) and speed_dev
 * (or NULL) as in sc_rk_stage_ex, for the z sweep of an RK form.  No
 * reference counterpart. */
sc_status sc_sweep_only_timed(sc_ctx* ctx, int dim, int form, float c0, float c1,
                              const float* dt_dev, const float* const* u0,
                              float* const* uk, float* const* out, int flags,
                              float* speed_dev, void* ev_start, void* ev_stop,
                              void* stream);

/* Launch geometry of the sweep kernels (no reference counterpart: the
 * reference leaves scheduling to XLA).  seg_*: cells per thread of a marching
 * sweep (>= 4); by_*: 64-lane waves per block (1..4).  Chosen per grid by
 * swirlc_amd.core.autotune, or set by the caller; results do not depend on
 * it (bit for bit). */
typedef struct sc_tuning {
  int seg_x, by_x;
  int seg_z, by_z;
  int by_y;
  int xcd_remap;            /* 1: XCD-aware block order (single-grid sweeps) */
  /* Adjoint passes (sc_rhs_vjp, sc_rk_stage_vjp): cells per thread and
   * waves per block of the marching x / z passes, waves per block of the y
   * pass.  Ignored by batches. */
  int adj_seg, adj_by;
  int adj_by_y;
  /* y sweep: z planes per wave (1..2^20): a wave keeps its 64-lane window
   * and marches over seg_y planes, loading each next plane's operands while
   * the current one is computed. */
  int seg_y;
} sc_tuning;

sc_status sc_ctx_get_tuning(const sc_ctx* ctx, sc_tuning* out);
sc_status sc_ctx_set_tuning(sc_ctx* ctx, const sc_tuning* tuning);

/* ---- Adjoint (no reference counterpart) ------------------------------------
 * The reference differentiates its TensorFlow graph with tf.GradientTape;
 * these entry points are hand-written reverse-mode counterparts of sc_rhs and
 * sc_rk_stage, for gradients of losses of RK rollouts.  Every face's forward
 * arithmetic is recomputed and differentiated exactly as executed (HLLC
 * branch, wave-speed min/max and MUSCL limiter selections replayed; WENO
 * smoothness-indicator and weight derivatives included).  dt is a constant
 * (no derivative through a CFL time step).  Boundary face / flux values and
 * a reference state (sc_ctx_set_reference_state) and primitive-variable BCs
 * are differentiated too, and one or two transported scalars (without
 * boundary values or a reference state).  p_ref and rho_ref are constants
 * of the derivative.  Cotangents ("lam") are full-size field arrays of
 * which only interior cells are read; outputs have zero halo cells. */

/* Allocates the context's adjoint work arrays (14 full-size fp32 fields);
 * required once before the calls below, which allocate nothing. */
sc_status sc_ctx_enable_adjoint(sc_ctx* ctx);

/* lq = (d rhs / d q)^T g: the vector-Jacobian product of sc_rhs at q
 * (equations/compressible_navier_stokes.py:69-268 differentiated, halo fill
 * included, so lq is the adjoint of q's interior cells).  Fills q's halos in
 * place. */
sc_status sc_rhs_vjp(sc_ctx* ctx, float* const* q, const float* const* g,
                     float* const* lq, void* stream);

/* Reverse of one sc_rk_stage (form SC_RK_EULER or SC_RK_BLEND) given the
 * adjoint lam_out of its output:
 *   lam_uk  = c1 * (lam_out + dt * (d rhs / d uk)^T lam_out)   (c1 = 1 for EULER)
 *   lam_u0 += c0 * lam_out                                     (BLEND, if not NULL)
 * lam_uk may alias lam_out.  Fills uk's halos in place. */
sc_status sc_rk_stage_vjp(sc_ctx* ctx, int form, float c0, float c1,
                          const float* dt_dev, float* const* uk,
                          const float* const* lam_out, float* const* lam_uk,
                          float* const* lam_u0, void* stream);

/* ---- Batched ensembles (no reference counterpart) ------------------------
 * Many independent grids of different sizes sharing one configuration,
 * advanced together: each phase (halo fill per dim, x / y / z sweep, wave
 * speed) is one launch over a packed list of (instance, block) tiles, with
 * per-instance time steps read from a device array, so a whole RK step of
 * the ensemble is a fixed launch sequence that can be captured in a HIP graph.
 * Instance i's field f occupies elements [offset, offset + (nx+2h)(ny+2h)(nz+2h))
 * of arena f, laid out like a single-grid field.  5 to 7 fields (up to two
 * transported scalars); gravity (each instance's potential from its own
 * coordinates), primitive-variable BCs and scalar face / flux boundary
 * values are supported; per-layer Dirichlet planes and plane-valued face /
 * flux values are not (a plane has the shape of one grid). */
typedef struct sc_instance {
  int nx, ny, nz;
  double dx, dy, dz;
  long long offset;
  /* Host cell-centre coordinates including halo cells (nx + 2h, ny + 2h,
   * nz + 2h entries, as sc_grid's), read at creation; required when the
   * parameters carry a gravity vector, else ignored. */
  const float* x_full;
  const float* y_full;
  const float* z_full;
} sc_instance;

typedef struct sc_batch sc_batch;

sc_status sc_batch_create(const sc_params* params, int halo, int n_instances,
                          const sc_instance* instances, sc_batch** out);
sc_status sc_batch_destroy(sc_batch* batch);
/* sc_fill_halos for every instance (arena base pointers). */
sc_status sc_batch_fill_halos(sc_batch* batch, float* const* q, void* stream);
/* sc_rk_stage for every instance; dt_dev holds one float per instance. */
sc_status sc_batch_rk_stage(sc_batch* batch, int form, float c0, float c1,
                            const float* dt_dev, const float* const* u0,
                            float* const* uk, float* const* out, void* stream);
/* sc_rk_stage_ex's flags for every instance (no reference counterpart):
 * SC_STAGE_HALOS_READY skips the fill of uk; SC_STAGE_FILL_OUT leaves out's
 * halos filled.  Where the fill is a pure copy along every dim - each dim
 * PERIODIC on both faces (every instance's extent >= 2h) or NEUMANN with
 * value 0 on both faces, the same for every field, no BC on a primitive -
 * the z sweep writes each output cell's images, the halo cells the fill
 * would copy from it (the wrapped or the nearest interior cell along each
 * dim); otherwise sc_batch_fill_halos runs on out.  speed_dev (or NULL): one
 * float per instance in device memory into which the z sweep max-accumulates
 * the wave speed of the instance's output cells (as sc_batch_max_wave_speed
 * computes it, bit for bit; clear it with sc_batch_cfl_dt or to 0 first). */
sc_status sc_batch_rk_stage_ex(sc_batch* batch, int form, float c0, float c1,
                               const float* dt_dev, const float* const* u0,
                               float* const* uk, float* const* out, int flags,
                               float* speed_dev, void* stream);
/* sc_cfl_dt per instance: dt_dev[i] = cfl * min(dx_i, dy_i, dz_i) /
 * speed_dev[i], then speed_dev[i] cleared to 0.  No reference counterpart. */
sc_status sc_batch_cfl_dt(sc_batch* batch, float* speed_dev, float cfl, float* dt_dev,
                          void* stream);
/* Tuning of the batched sweeps (sc_tuning above).  Setting it rebuilds the
 * tile lists (allocates; not allowed during graph capture). */
sc_status sc_batch_get_tuning(const sc_batch* batch, sc_tuning* out);
sc_status sc_batch_set_tuning(sc_batch* batch, const sc_tuning* tuning);
/* Per-instance maximum wave speed (speed_dev[i]) and, if dt_dev is not NULL,
 * dt_dev[i] = cfl * min(dx_i, dy_i, dz_i) / speed_dev[i]. */
sc_status sc_batch_max_wave_speed(sc_batch* batch, const float* const* q, float cfl,
                                  float* speed_dev, float* dt_dev, void* stream);

/* Thread-local message of the last failed call ("" if none). */
const char* sc_last_error(void);

/* Library version string. */
const char* sc_version(void);

#ifdef __cplusplus
}
#endif

#endif /* SWIRLC_H_ */
