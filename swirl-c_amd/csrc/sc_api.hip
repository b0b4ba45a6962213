// C ABI of libswirlc_amd (declared in include/swirlc.h).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <string>

#include "../../include/swirlc.h"
#include "sc_adjoint_impl.hpp"
#include "sc_halo.hpp"
#include "sc_diffusion.hpp"
#include "sc_interp.hpp"
#include "sc_riemann.hpp"
#include "sc_launch.hpp"

struct sc_ctx {
  sc_grid grid;
  sc_params params;
  int device;
  int nf;
  int N[3];
  long long cells_full;
  float h[3];
  // gravitational potential tables (device): pe = (pe_x[i] - pe_y[j]) - pe_z[k]
  int has_pe;
  float* pe_x;
  float* pe_y;
  float* pe_z;
  int grav_mask;
  // scratch RHS accumulator (device), one full-size array per field
  float* rhs_scratch[SC_MAX_FIELDS];
  unsigned* speed_bits;
  // boundary face / flux values per [dim][side] (device) and dims using them
  sc::BoundaryBc* bbc_dev;
  int bbc_dims;
  // adjoint work arrays (sc_ctx_enable_adjoint): state adjoint and the nine
  // velocity-gradient adjoints
  float* adj_lq[SC_MAX_FIELDS];
  float* adj_lG[9];
  sc::HaloArgs halo;  // BC tables; field pointers filled per call
  int halo_all_periodic;  // one-launch periodic fill applies
  // A stage's last sweep writes out's halos as images (all-periodic grids:
  // copy_fill_dims with no clamping dim); img_clamp as in SweepArgs.
  int images_ok;
  int img_clamp;
  sc::SweepArgs sweep;
  sc::SweepTuning tune;
};

// RHS accumulator of sc_rk_stage_noalloc: a file-scope device array, one copy
// per GPU, held until the process ends.  Host code never passes `scratch`
// itself to a kernel (on the host it is only a stand-in for the symbol): the
// device address of each GPU's copy is looked up once per device with
// hipGetSymbolAddress (at context creation, with that device current) and
// kept in g_scratch_ptr.
constexpr long long kScratchFloats = 32LL << 20;  // 128 MiB
static __device__ __align__(256) float scratch[kScratchFloats];

namespace {

thread_local std::string g_last_error;

constexpr int kMaxDevices = 64;
float* g_scratch_ptr[kMaxDevices];
std::mutex g_scratch_mu;

sc_status fail(sc_status code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

sc_status hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return SC_OK;
  return fail(SC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Whether the halo fill of every field is a pure copy along every dim, so
// that a halo cell equals the interior cell at its wrapped (periodic: both
// faces PERIODIC, extent >= 2h) or clamped (both faces NEUMANN with value
// +0: ghost = neighbour - 0, cumulatively, copies the edge cell bit for bit)
// index in each dim; every field must take the same kind along a dim, and no
// BC may be given on a primitive (those fill in primitive space).  Sets bit d
// of *clamp for the clamping dims (d = 0 x, 1 y, 2 z).
bool copy_fill_dims(const sc_params& p, int nf, const int n[3], int h, int* clamp) {
  if (p.prim_bc_mask != 0) return false;
  int cl = 0;
  for (int d = 0; d < 3; ++d) {
    const int kind = p.bc[0][d][0].type;
    for (int f = 0; f < nf; ++f)
      for (int s = 0; s < 2; ++s) {
        const sc_bc_face& b = p.bc[f][d][s];
        if (b.type != kind) return false;
        if (kind == SC_BC_NEUMANN &&
            (b.value != 0.0f || signbit(b.value) || b.n_layer_values != 0 ||
             b.layer_planes != nullptr))
          return false;
      }
    if (kind == SC_BC_PERIODIC) {
      if (n[d] < 2 * h) return false;
    } else if (kind == SC_BC_NEUMANN) {
      cl |= 1 << d;
    } else {
      return false;
    }
  }
  *clamp = cl;
  return true;
}

// common/utils.py:114-142.
int gravity_direction(const double g[3]) {
  const double norm = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
  const double small = 1e-6;  // finfo(float32).resolution
  if (fabs(norm - 1.0) >= small) return -1;
  for (int i = 0; i < 3; ++i)
    if (fabs(fabs(g[i]) - 1.0) < small) return i;
  return 3;
}

int radius_of(int scheme) { return scheme == SC_WENO_5 ? 2 : 1; }

void fill_halo_fields(sc_ctx* c, float* const* q) {
  for (int f = 0; f < c->nf; ++f) c->halo.q[f] = q[f];
}

void fill_sweep_state(sc_ctx* c, const float* const* q) {
  for (int f = 0; f < c->nf; ++f) c->sweep.q.p[f] = q[f];
}

// Blocks are 64 x by threads and the kernels are built for at most 256.
sc_status to_tuning(const sc_tuning* t, sc::SweepTuning* out) {
  if (t == nullptr) return fail(SC_ERR_INVALID, "null tuning");
  const int seg[3] = {t->seg_x, t->seg_z, t->adj_seg};
  const int by[5] = {t->by_x, t->by_z, t->by_y, t->adj_by, t->adj_by_y};
  for (int v : seg)
    if (v < 4 || v > (1 << 20)) return fail(SC_ERR_INVALID, "tuning: seg_x, seg_z, adj_seg must be in [4, 2^20]");
  for (int v : by)
    if (v < 1 || v > 4) return fail(SC_ERR_INVALID, "tuning: by_x, by_z, by_y, adj_by, adj_by_y must be in [1, 4]");
  if (t->xcd_remap != 0 && t->xcd_remap != 1)
    return fail(SC_ERR_INVALID, "tuning: xcd_remap must be 0 or 1");
  if (t->seg_y < 1 || t->seg_y > (1 << 20))
    return fail(SC_ERR_INVALID, "tuning: seg_y must be in [1, 2^20]");
  out->seg_x = t->seg_x;
  out->by_x = t->by_x;
  out->seg_z = t->seg_z;
  out->by_z = t->by_z;
  out->by_y = t->by_y;
  out->xcd_remap = t->xcd_remap;
  out->adj_seg = t->adj_seg;
  out->adj_by = t->adj_by;
  out->adj_by_y = t->adj_by_y;
  out->seg_y = t->seg_y;
  return SC_OK;
}

// Per-(dim, side) device tables of the boundary face / flux values.
sc_status build_boundary_tables(const sc_params& p, sc::BoundaryBc (*t)[2], int* dims) {
  memset(t, 0, sizeof(sc::BoundaryBc) * 6);
  *dims = 0;
  if (p.n_boundary_values < 0 || p.n_boundary_values > SC_MAX_BOUNDARY_VALUES)
    return fail(SC_ERR_INVALID, "n_boundary_values out of range");
  for (int k = 0; k < p.n_boundary_values; ++k) {
    const sc_boundary_value& v = p.boundary_values[k];
    if (v.dim < 0 || v.dim > 2 || v.side < 0 || v.side > 1)
      return fail(SC_ERR_INVALID, "boundary value: dim must be 0..2 and side 0..1");
    sc::BoundaryBc& b = t[v.dim][v.side];
    if (v.kind >= SC_BV_FLUX_CONVECTIVE && v.kind <= SC_BV_FLUX_TOTAL) {
      if (v.var < 0 || v.var >= p.n_fields)
        return fail(SC_ERR_INVALID, "flux boundary value: var must be a conserved field index");
      b.flux_mask |= 1 << (8 * v.kind + v.var);
      b.flux_val[v.kind][v.var] = v.value;
      b.flux_plane[v.kind][v.var] = v.plane;
    } else if (v.kind == SC_BV_FACE_CONSERVED || v.kind == SC_BV_FACE_PRIMITIVE) {
      if (p.include_diffusion)
        return fail(SC_ERR_UNSUPPORTED,
                    "Face boundary conditions are not supported with diffusion enabled at"
                    " this time.");
      int var = v.var;
      if (v.kind == SC_BV_FACE_CONSERVED) {
        if (var == 0)
          return fail(SC_ERR_UNSUPPORTED,
                      "a density face value replaces the face primitives without p and h"
                      " (boundary.py:284-293), which the flux then lacks");
        if (var < 1 || var >= p.n_fields)
          return fail(SC_ERR_INVALID, "face boundary value: invalid conserved field index");
        if (var >= 5) var = 7 + (var - 5);  // scalar primitive slot
      } else if (var < 1 || var >= 7 + (p.n_fields - 5)) {
        return fail(SC_ERR_INVALID, "face boundary value: invalid primitive index");
      }
      if (b.n_face >= SC_MAX_FACE_VALUES)
        return fail(SC_ERR_INVALID, "too many face boundary values on one boundary");
      b.face[b.n_face++] = {v.kind == SC_BV_FACE_CONSERVED ? 1 : 2, var, v.value, v.plane};
    } else {
      return fail(SC_ERR_INVALID, "boundary value: unknown kind");
    }
    *dims |= 1 << v.dim;
  }
  return SC_OK;
}

void from_tuning(const sc::SweepTuning& t, sc_tuning* out) {
  out->seg_x = t.seg_x;
  out->by_x = t.by_x;
  out->seg_z = t.seg_z;
  out->by_z = t.by_z;
  out->by_y = t.by_y;
  out->xcd_remap = t.xcd_remap;
  out->adj_seg = t.adj_seg;
  out->adj_by = t.adj_by;
  out->adj_by_y = t.adj_by_y;
  out->seg_y = t.seg_y;
}

}  // namespace

extern "C" {

const char* sc_last_error(void) { return g_last_error.c_str(); }

const char* sc_version(void) { return "swirlc_amd 0.1 gfx950"; }

sc_status sc_ctx_create(const sc_grid* grid, const sc_params* params, sc_ctx** out) {
  if (grid == nullptr || params == nullptr || out == nullptr)
    return fail(SC_ERR_INVALID, "sc_ctx_create: null argument");
  *out = nullptr;
  const sc_params& p = *params;
  if (p.interpolation < SC_WENO_3 || p.interpolation > SC_MUSCL)
    return fail(SC_ERR_UNSUPPORTED,
                "interpolation_scheme is not implemented. Valid options are: "
                "['WENO_3', 'WENO_5', 'MUSCL']");
  if (p.numeric_flux != SC_HLL && p.numeric_flux != SC_HLLC)
    return fail(SC_ERR_INVALID,
                "numeric_flux_scheme is not implemented. Valid options are: "
                "['HLL', 'HLLC']");
  if (p.n_fields < 5 || p.n_fields > SC_MAX_FIELDS)
    return fail(SC_ERR_UNSUPPORTED, "n_fields must be 5, 6 or 7 (at most two transported scalars)");
  if (grid->nx < 1 || grid->ny < 1 || grid->nz < 1)
    return fail(SC_ERR_INVALID, "grid extents must be positive");
  const int h = grid->halo;
  if (h < radius_of(p.interpolation) + 1 || h > SC_MAX_HALO)
    return fail(SC_ERR_INVALID, "halo_width too small for the interpolation stencil"
                                " (WENO_5 needs 3, WENO_3/MUSCL need 2) or above 4");
  if (grid->nx >= (1 << 20) ||
      (long long)(grid->nx + 2 * h) * (grid->ny + 2 * h) >= (1LL << 30))
    return fail(SC_ERR_UNSUPPORTED, "grid too large: nx must be below 2^20 and "
                                    "(nx + 2 * halo_width) * (ny + 2 * halo_width) below 2^30");
  if (!(grid->dx > 0 && grid->dy > 0 && grid->dz > 0))
    return fail(SC_ERR_INVALID, "grid spacing must be positive");
  const int gdir = gravity_direction(p.g);
  if (p.gravity_source && gdir == -1)
    return fail(SC_ERR_INVALID,
                "Gravity direction must be specified if the gravity source function"
                " is to be used.");
  const int n[3] = {grid->nx, grid->ny, grid->nz};
  for (int f = 0; f < p.n_fields; ++f) {
    for (int d = 0; d < 3; ++d) {
      for (int s = 0; s < 2; ++s) {
        const sc_bc_face& b = p.bc[f][d][s];
        if (b.type == SC_BC_REFLECTIVE)
          return fail(SC_ERR_UNSUPPORTED, "Unsupported BC type BoundaryCondition.REFLECTIVE.");
        if (b.type < SC_BC_NONE || b.type > SC_BC_NEUMANN)
          return fail(SC_ERR_INVALID, "unknown BC type");
        if ((b.type == SC_BC_NONE) != (p.bc[f][d][1 - s].type == SC_BC_NONE))
          return fail(SC_ERR_INVALID, "SC_BC_NONE must be given on both faces of a dim");
        if (b.type == SC_BC_NONE && ((p.prim_bc_mask >> f) & 1))
          return fail(SC_ERR_UNSUPPORTED, "a BC given on a primitive must cover every dim");
        if (b.type == SC_BC_PERIODIC && n[d] < h)
          return fail(SC_ERR_UNSUPPORTED, "PERIODIC BC needs at least halo_width interior cells");
        if (b.n_layer_values != 0 && b.n_layer_values != h)
          return fail(SC_ERR_INVALID, "per-layer DIRICHLET values must have halo_width entries");
      }
    }
  }
  if (p.prim_bc_mask & 1)
    return fail(SC_ERR_INVALID, "the density BC cannot be given on a primitive");
  sc::BoundaryBc bbc[3][2];
  int bbc_dims = 0;
  {
    const sc_status st = build_boundary_tables(p, bbc, &bbc_dims);
    if (st != SC_OK) return st;
  }
  if (gdir != -1 && (grid->x_full == nullptr || grid->y_full == nullptr || grid->z_full == nullptr))
    return fail(SC_ERR_INVALID, "coordinates are required when the gravity vector is set");

  sc_ctx* c = new sc_ctx();
  memset(c, 0, sizeof(sc_ctx));
  c->tune = sc::SweepTuning();
  c->grid = *grid;
  c->params = p;
  c->nf = p.n_fields;
  hipGetDevice(&c->device);
  {
    // Device address of this GPU's no-allocation scratch array (once per device).
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (c->device >= 0 && c->device < kMaxDevices && g_scratch_ptr[c->device] == nullptr) {
      float* ptr = nullptr;
      const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
      if (e != hipSuccess) {
        delete c;
        return hip_check(e, "hipGetSymbolAddress(scratch)");
      }
      g_scratch_ptr[c->device] = ptr;
    }
  }
  for (int d = 0; d < 3; ++d) c->N[d] = n[d] + 2 * h;
  c->cells_full = (long long)c->N[0] * c->N[1] * c->N[2];
  c->h[0] = (float)grid->dx;
  c->h[1] = (float)grid->dy;
  c->h[2] = (float)grid->dz;

  sc_status st = SC_OK;
  auto check = [&](hipError_t e, const char* what) {
    if (st == SC_OK) st = hip_check(e, what);
  };
  c->has_pe = gdir != -1;
  if (c->has_pe) {
    // pe = -G g . x as ((-G gx) * x - (G gy) * y) - (G gz) * z in fp32
    // (generic.py:312-319): tabulate each term per axis.
    const float* coord[3] = {grid->x_full, grid->y_full, grid->z_full};
    float** dst[3] = {&c->pe_x, &c->pe_y, &c->pe_z};
    for (int d = 0; d < 3; ++d) {
      const double gv = 9.81 * p.g[d];
      const float coef = d == 0 ? (float)(-gv) : (float)gv;
      float* tmp = (float*)malloc(sizeof(float) * c->N[d]);
      for (int i = 0; i < c->N[d]; ++i) tmp[i] = coef * coord[d][i];
      check(hipMalloc((void**)dst[d], sizeof(float) * c->N[d]), "hipMalloc pe");
      if (st == SC_OK)
        check(hipMemcpy(*dst[d], tmp, sizeof(float) * c->N[d], hipMemcpyHostToDevice),
              "hipMemcpy pe");
      free(tmp);
    }
  }
  c->grav_mask = 0;
  if (p.gravity_source) c->grav_mask = gdir == 3 ? 7 : (1 << gdir);
  for (int f = 0; f < c->nf && st == SC_OK; ++f) {
    check(hipMalloc((void**)&c->rhs_scratch[f], sizeof(float) * c->cells_full), "hipMalloc rhs");
  }
  if (st == SC_OK) check(hipMalloc((void**)&c->speed_bits, sizeof(unsigned) * 4), "hipMalloc speed");
  c->bbc_dims = bbc_dims;
  if (st == SC_OK && bbc_dims) {
    check(hipMalloc((void**)&c->bbc_dev, sizeof(bbc)), "hipMalloc boundary values");
    if (st == SC_OK)
      check(hipMemcpy(c->bbc_dev, bbc, sizeof(bbc), hipMemcpyHostToDevice),
            "hipMemcpy boundary values");
  }
  if (st == SC_OK) {
    for (int f = 0; f < c->nf; ++f)
      check(hipMemset(c->rhs_scratch[f], 0, sizeof(float) * c->cells_full), "hipMemset rhs");
  }
  if (st != SC_OK) {
    sc_ctx_destroy(c);
    return st;
  }

  // Halo tables.
  sc::HaloArgs& ha = c->halo;
  ha.nf = c->nf;
  ha.h = h;
  for (int d = 0; d < 3; ++d) {
    ha.n[d] = n[d];
    ha.N[d] = c->N[d];
  }
  ha.st[0] = c->N[1];
  ha.st[1] = 1;
  ha.st[2] = (long long)c->N[0] * c->N[1];
  ha.prim_mask = p.prim_bc_mask;
  for (int f = 0; f < c->nf; ++f)
    for (int d = 0; d < 3; ++d)
      for (int s = 0; s < 2; ++s) {
        const sc_bc_face& b = p.bc[f][d][s];
        sc::BcFaceDev& o = ha.bc[f][d][s];
        o.type = b.type;
        o.value = b.value;
        o.n_layer_values = b.n_layer_values;
        for (int i = 0; i < SC_MAX_HALO; ++i) o.layer_values[i] = b.layer_values[i];
        o.layer_planes = b.layer_planes;
      }
  c->halo_all_periodic = ha.prim_mask == 0 && sc::halo_periodic_threads(ha) < (1LL << 31);
  for (int f = 0; f < c->nf; ++f)
    for (int d = 0; d < 3; ++d)
      for (int s = 0; s < 2; ++s)
        if (p.bc[f][d][s].type != SC_BC_PERIODIC) c->halo_all_periodic = 0;
  {
    // Single grids take the images on all-periodic grids only: a march whose
    // edge cells write h clamped images each (NEUMANN 0) was measured slower
    // than the fill after the sweeps on grids of 32^3 to 160x64x64 (z sweep
    // 57 -> 123 us: the edge waves' extra stores lengthen a launch with few
    // waves), while the batched launches, with many waves, gain.
    const int n3[3] = {c->grid.nx, c->grid.ny, c->grid.nz};
    c->img_clamp = 0;
    c->images_ok = copy_fill_dims(p, c->nf, n3, c->grid.halo, &c->img_clamp) && c->img_clamp == 0;
    c->img_clamp = 0;  // the single-grid sweeps are built without clamped images
  }

  // Sweep constants.
  sc::SweepArgs& a = c->sweep;
  a.nx = n[0];
  a.ny = n[1];
  a.nz = n[2];
  a.h = h;
  a.NX = c->N[0];
  a.NY = c->N[1];
  a.NZ = c->N[2];
  a.sx = c->N[1];
  a.sz = (long long)c->N[0] * c->N[1];
  a.has_pe = c->has_pe;
  a.pe_x = c->pe_x;
  a.pe_y = c->pe_y;
  a.pe_z = c->pe_z;
  a.grav_src = p.gravity_source ? 1 : 0;
  a.grav_mask = c->grav_mask;
  for (int d = 0; d < 3; ++d) a.gdir[d] = (float)p.g[d];
  a.prim_bc_mask = p.prim_bc_mask;
  a.vc.nu = (float)p.nu;
  // kappa = ((1/Pr) * cp) * mu with (1/Pr) and cp as fp32 (simple.py:104-108).
  a.vc.kfac = (float)(1.0 / p.pr) * sc::kCp;
  for (int d = 0; d < 3; ++d) {
    a.vc.h[d] = c->h[d];
    a.vc.inv_h[d] = 1.0f / c->h[d];
  }
  a.bbc_dims = c->bbc_dims;
  a.bbc = c->bbc_dev;
  a.vc.inv_two_h[0] = 1.0f / (float)(2.0 * grid->dx);
  a.vc.inv_two_h[1] = 1.0f / (float)(2.0 * grid->dy);
  a.vc.inv_two_h[2] = 1.0f / (float)(2.0 * grid->dz);
  // Coordinates are not kept: the context only needs the pe tables.
  c->grid.x_full = c->grid.y_full = c->grid.z_full = nullptr;
  *out = c;
  return SC_OK;
}

sc_status sc_ctx_destroy(sc_ctx* c) {
  if (c == nullptr) return SC_OK;
  if (c->pe_x) hipFree(c->pe_x);
  if (c->pe_y) hipFree(c->pe_y);
  if (c->pe_z) hipFree(c->pe_z);
  for (int f = 0; f < SC_MAX_FIELDS; ++f)
    if (c->rhs_scratch[f]) hipFree(c->rhs_scratch[f]);
  if (c->speed_bits) hipFree(c->speed_bits);
  if (c->bbc_dev) hipFree(c->bbc_dev);
  for (int f = 0; f < SC_MAX_FIELDS; ++f)
    if (c->adj_lq[f]) hipFree(c->adj_lq[f]);
  for (int f = 0; f < 9; ++f)
    if (c->adj_lG[f]) hipFree(c->adj_lG[f]);
  delete c;
  return SC_OK;
}

sc_status sc_fill_halos(sc_ctx* c, float* const* q, void* stream) {
  if (c == nullptr || q == nullptr) return fail(SC_ERR_INVALID, "sc_fill_halos: null argument");
  for (int f = 0; f < c->nf; ++f)
    if (q[f] == nullptr) return fail(SC_ERR_INVALID, "sc_fill_halos: null field pointer");
  hipStream_t s = (hipStream_t)stream;
  fill_halo_fields(c, q);
  if (c->halo_all_periodic) return hip_check(sc::launch_halo_periodic(c->halo, s), "halo kernel");
  for (int d = 0; d < 3; ++d) {
    sc_status st = hip_check(sc::launch_halo_dim(c->halo, d, s), "halo kernel");
    if (st != SC_OK) return st;
  }
  if (c->halo.prim_mask)
    return hip_check(sc::launch_halo_prim_finish(c->halo, s), "halo prim kernel");
  return SC_OK;
}

static sc_status run_sweeps(sc_ctx* c, const sc::SweepArgs& a, hipStream_t s,
                            int only_dir = -1, const sc::SweepTuning* tune = nullptr) {
  const sc::SweepTuning& t = tune != nullptr ? *tune : c->tune;
  hipError_t e;
  const int scheme = c->params.interpolation;
  const int flux = c->params.numeric_flux;
  const bool visc = c->params.include_diffusion != 0;
  const int ns = c->nf - 5;
  if (scheme == SC_WENO_5) {
    e = sc::launch_rhs_sweeps<sc::kWeno5>(a, c->h, flux, visc, ns, t, s, only_dir);
  } else if (scheme == SC_WENO_3) {
    e = sc::launch_rhs_sweeps<sc::kWeno3>(a, c->h, flux, visc, ns, t, s, only_dir);
  } else {
    e = sc::launch_rhs_sweeps<sc::kMuscl>(a, c->h, flux, visc, ns, t, s, only_dir);
  }
  return hip_check(e, "sweep kernels");
}

sc_status sc_rhs(sc_ctx* c, float* const* q, float* const* rhs, void* stream) {
  if (c == nullptr || q == nullptr || rhs == nullptr)
    return fail(SC_ERR_INVALID, "sc_rhs: null argument");
  for (int f = 0; f < c->nf; ++f)
    if (q[f] == nullptr || rhs[f] == nullptr)
      return fail(SC_ERR_INVALID, "sc_rhs: null field pointer");
  sc_status st = sc_fill_halos(c, q, stream);
  if (st != SC_OK) return st;
  sc::SweepArgs a = c->sweep;
  for (int f = 0; f < c->nf; ++f) {
    a.q.p[f] = q[f];
    a.r.p[f] = rhs[f];
  }
  a.rk_form = 0;
  return run_sweeps(c, a, (hipStream_t)stream);
}

static sc_status rk_stage_with(sc_ctx* c, int form, float c0, float c1,
                               const float* dt_dev, const float* const* u0,
                               float* const* uk, float* const* out, float* const* racc,
                               int flags, unsigned* speed_bits, void* stream);

sc_status sc_rk_stage(sc_ctx* c, int form, float c0, float c1, const float* dt_dev,
                      const float* const* u0, float* const* uk, float* const* out,
                      void* stream) {
  if (c == nullptr) return fail(SC_ERR_INVALID, "sc_rk_stage: null argument");
  return rk_stage_with(c, form, c0, c1, dt_dev, u0, uk, out, c->rhs_scratch, 0, nullptr,
                       stream);
}

sc_status sc_rk_stage_ex(sc_ctx* c, int form, float c0, float c1, const float* dt_dev,
                         const float* const* u0, float* const* uk, float* const* out, int flags,
                         float* speed_dev, void* stream) {
  if (c == nullptr) return fail(SC_ERR_INVALID, "sc_rk_stage_ex: null argument");
  if (flags & ~(SC_STAGE_HALOS_READY | SC_STAGE_FILL_OUT))
    return fail(SC_ERR_INVALID, "sc_rk_stage_ex: unknown flag");
  return rk_stage_with(c, form, c0, c1, dt_dev, u0, uk, out, c->rhs_scratch, flags,
                       (unsigned*)speed_dev, stream);
}

static float min_spacing(const sc_ctx* c) {
  float hmin = c->h[0];
  if (c->h[1] < hmin) hmin = c->h[1];
  if (c->h[2] < hmin) hmin = c->h[2];
  return hmin;
}

sc_status sc_cfl_dt(sc_ctx* c, float* speed_dev, float cfl, float* dt_dev, void* stream) {
  if (c == nullptr || speed_dev == nullptr || dt_dev == nullptr)
    return fail(SC_ERR_INVALID, "sc_cfl_dt: null argument");
  return hip_check(sc::launch_cfl_dt_reset((unsigned*)speed_dev, cfl, min_spacing(c), dt_dev,
                                           (hipStream_t)stream),
                   "cfl dt kernel");
}

sc_status sc_rk_stage_noalloc(sc_ctx* c, int form, float c0, float c1, const float* dt_dev,
                              const float* const* u0, float* const* uk, float* const* out,
                              void* stream) {
  if (c == nullptr) return fail(SC_ERR_INVALID, "sc_rk_stage_noalloc: null argument");
  // Each field's accumulator starts on a 256-byte boundary.
  const long long stride = (c->cells_full + 63) / 64 * 64;
  if (stride * c->nf > kScratchFloats)
    return fail(SC_ERR_UNSUPPORTED,
                "sc_rk_stage_noalloc: grid too large for the 128 MiB scratch array (needs"
                " n_fields * (nx+2h)(ny+2h)(nz+2h), each rounded up to 64, <= 33554432"
                " floats); use sc_rk_stage");
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev != c->device)
    return fail(SC_ERR_INVALID, "sc_rk_stage_noalloc: the context's device must be current");
  float* base = g_scratch_ptr[c->device];
  if (base == nullptr)
    return fail(SC_ERR_HIP, "sc_rk_stage_noalloc: scratch array address not resolved");
  float* racc[SC_MAX_FIELDS] = {};
  for (int f = 0; f < c->nf; ++f) racc[f] = base + f * stride;
  return rk_stage_with(c, form, c0, c1, dt_dev, u0, uk, out, racc, 0, nullptr, stream);
}

static sc_status rk_stage_with(sc_ctx* c, int form, float c0, float c1,
                               const float* dt_dev, const float* const* u0,
                               float* const* uk, float* const* out, float* const* racc,
                               int flags, unsigned* speed_bits, void* stream) {
  if (c == nullptr || uk == nullptr || out == nullptr || dt_dev == nullptr)
    return fail(SC_ERR_INVALID, "sc_rk_stage: null argument");
  if (form == SC_RK_AVG)
    return fail(SC_ERR_UNSUPPORTED,
                "sc_rk_stage: the SC_RK_AVG form is disabled (an RK2 run on the GPU ended"
                " in a device fault whose cause is not yet identified); evaluate RK2 with"
                " sc_rhs and element-wise updates");
  if (form != SC_RK_EULER && form != SC_RK_BLEND)
    return fail(SC_ERR_INVALID, "sc_rk_stage: unknown RK form");
  if (form != SC_RK_EULER && u0 == nullptr)
    return fail(SC_ERR_INVALID, "sc_rk_stage: u0 required for this RK form");
  for (int f = 0; f < c->nf; ++f) {
    if (uk[f] == nullptr || out[f] == nullptr)
      return fail(SC_ERR_INVALID, "sc_rk_stage: null field pointer");
    if (out[f] == uk[f]) return fail(SC_ERR_INVALID, "sc_rk_stage: out must not alias uk");
    if (form != SC_RK_EULER && u0[f] == nullptr)
      return fail(SC_ERR_INVALID, "sc_rk_stage: null u0 field pointer");
  }
  if (!(flags & SC_STAGE_HALOS_READY)) {
    sc_status st = sc_fill_halos(c, uk, stream);
    if (st != SC_OK) return st;
  }
  sc::SweepArgs a = c->sweep;
  for (int f = 0; f < c->nf; ++f) {
    a.q.p[f] = uk[f];
    a.u0.p[f] = form == SC_RK_EULER ? nullptr : u0[f];
    a.r.p[f] = racc[f];
    a.out.p[f] = out[f];
  }
  a.rk_form = form;
  a.c0 = c0;
  a.c1 = c1;
  a.dt = dt_dev;
  a.speed_bits = speed_bits;
  // out's halos: written by the last sweep as images where that is exactly
  // the fill (copy_fill_dims), else filled after.
  const bool fill_out = (flags & SC_STAGE_FILL_OUT) != 0;
  a.out_images = fill_out && c->images_ok;
  a.img_clamp = c->img_clamp;
  sc_status st = run_sweeps(c, a, (hipStream_t)stream);
  if (st != SC_OK || !fill_out || a.out_images) return st;
  return sc_fill_halos(c, out, stream);
}

static sc_status sweep_only_with(sc_ctx* c, int dim, int form, float c0, float c1,
                                 const float* dt_dev, const float* const* u0,
                                 float* const* uk, float* const* out, int flags,
                                 unsigned* speed_bits, void* stream, void* ev_start,
                                 void* ev_stop);

sc_status sc_sweep_only(sc_ctx* c, int dim, int form, float c0, float c1,
                        const float* dt_dev, const float* const* u0, float* const* uk,
                        float* const* out, void* stream) {
  return sweep_only_with(c, dim, form, c0, c1, dt_dev, u0, uk, out, 0, nullptr, stream,
                         nullptr, nullptr);
}

sc_status sc_sweep_only_timed(sc_ctx* c, int dim, int form, float c0, float c1,
                              const float* dt_dev, const float* const* u0, float* const* uk,
                              float* const* out, int flags, float* speed_dev,
                              void* ev_start, void* ev_stop, void* stream) {
  if (ev_start == nullptr || ev_stop == nullptr)
    return fail(SC_ERR_INVALID, "sc_sweep_only_timed: null event");
  return sweep_only_with(c, dim, form, c0, c1, dt_dev, u0, uk, out, flags,
                         (unsigned*)speed_dev, stream, ev_start, ev_stop);
}

static sc_status sweep_only_with(sc_ctx* c, int dim, int form, float c0, float c1,
                                 const float* dt_dev, const float* const* u0,
                                 float* const* uk, float* const* out, int flags,
                                 unsigned* speed_bits, void* stream, void* ev_start,
                                 void* ev_stop) {
  if (c == nullptr || uk == nullptr) return fail(SC_ERR_INVALID, "sc_sweep_only: null argument");
  if (dim < 0 || dim > 2) return fail(SC_ERR_INVALID, "sc_sweep_only: dim must be 0..2");
  if (form != 0 && (out == nullptr || dt_dev == nullptr || (form != SC_RK_EULER && u0 == nullptr)))
    return fail(SC_ERR_INVALID, "sc_sweep_only: missing RK operands");
  sc::SweepArgs a = c->sweep;
  for (int f = 0; f < c->nf; ++f) {
    a.q.p[f] = uk[f];
    a.u0.p[f] = (form == 0 || form == SC_RK_EULER) ? nullptr : u0[f];
    a.r.p[f] = c->rhs_scratch[f];
    a.out.p[f] = form == 0 ? nullptr : out[f];
  }
  a.rk_form = form;
  a.c0 = c0;
  a.c1 = c1;
  a.dt = dt_dev;
  a.speed_bits = form == 0 ? nullptr : speed_bits;
  a.out_images = form != 0 && (flags & SC_STAGE_FILL_OUT) && c->images_ok;
  a.img_clamp = c->img_clamp;
  sc::SweepTuning t = c->tune;
  t.ev_start = (hipEvent_t)ev_start;
  t.ev_stop = (hipEvent_t)ev_stop;
  return run_sweeps(c, a, (hipStream_t)stream, dim, &t);
}

sc_status sc_max_wave_speed(sc_ctx* c, const float* const* q, float cfl, float* out_dev,
                            float* dt_dev, void* stream) {
  if (c == nullptr || q == nullptr || out_dev == nullptr)
    return fail(SC_ERR_INVALID, "sc_max_wave_speed: null argument");
  sc::SpeedArgs a;
  for (int f = 0; f < 5; ++f) a.q[f] = q[f];
  a.nx = c->sweep.nx;
  a.ny = c->sweep.ny;
  a.nz = c->sweep.nz;
  a.h = c->sweep.h;
  a.sx = c->sweep.sx;
  a.sz = c->sweep.sz;
  a.has_pe = c->has_pe;
  a.pe_x = c->pe_x;
  a.pe_y = c->pe_y;
  a.pe_z = c->pe_z;
  const float hmin = min_spacing(c);
  return hip_check(sc::launch_max_speed(a, (unsigned*)out_dev, cfl, hmin, dt_dev,
                                        (hipStream_t)stream),
                   "max speed kernel");
}

sc_status sc_face_interp(const float* v, const int* extents, int dim, int scheme, float* neg,
                         float* pos, void* stream) {
  if (v == nullptr || extents == nullptr || neg == nullptr || pos == nullptr)
    return fail(SC_ERR_INVALID, "sc_face_interp: null argument");
  if (neg == v || pos == v || neg == pos)
    return fail(SC_ERR_INVALID, "sc_face_interp: v, neg and pos must be distinct arrays");
  if (dim < 0 || dim > 2)
    return fail(SC_ERR_INVALID, "sc_face_interp: dim must be 0 (x), 1 (y) or 2 (z)");
  if (extents[0] < 1 || extents[1] < 1 || extents[2] < 1)
    return fail(SC_ERR_INVALID, "sc_face_interp: extents must be positive");
  if (scheme != SC_WENO_3 && scheme != SC_WENO_5 && scheme != SC_MUSCL)
    return fail(SC_ERR_UNSUPPORTED,
                "interpolation_scheme is not implemented. Valid options are: "
                "['WENO_3', 'WENO_5', 'MUSCL']");
  const int s = scheme == SC_WENO_5 ? sc::kWeno5 : (scheme == SC_WENO_3 ? sc::kWeno3 : sc::kMuscl);
  return hip_check(sc::launch_face_interp(v, extents, dim, s, neg, pos, (hipStream_t)stream),
                   "face interpolation kernel");
}

sc_status sc_face_bc_elementwise(int op, const float* a, const float* b, float* out,
                                 long long n, void* stream) {
  if (a == nullptr || b == nullptr || out == nullptr)
    return fail(SC_ERR_INVALID, "sc_face_bc_elementwise: null argument");
  if (op < 0 || op > 2) return fail(SC_ERR_INVALID, "sc_face_bc_elementwise: op must be 0, 1 or 2");
  if (n < 1) return fail(SC_ERR_INVALID, "sc_face_bc_elementwise: n must be positive");
  return hip_check(sc::launch_elementwise(op, a, b, out, n, (hipStream_t)stream),
                   "elementwise kernel");
}

static sc_status fill_plane_with(const char* name, float* f, const int* extents, int dim,
                                 int index, float value, const float* values, void* stream) {
  if (f == nullptr || extents == nullptr)
    return fail(SC_ERR_INVALID, std::string(name) + ": null argument");
  if (dim < 0 || dim > 2)
    return fail(SC_ERR_INVALID, std::string(name) + ": dim must be 0 (x), 1 (y) or 2 (z)");
  if (extents[0] < 1 || extents[1] < 1 || extents[2] < 1)
    return fail(SC_ERR_INVALID, std::string(name) + ": extents must be positive");
  if (index < 0 || index >= extents[dim])
    return fail(SC_ERR_INVALID, std::string(name) + ": plane index outside the array");
  return hip_check(sc::launch_fill_plane(f, extents, dim, index, value, values,
                                         (hipStream_t)stream),
                   "plane fill kernel");
}

sc_status sc_fill_plane(float* f, const int* extents, int dim, int index, float value,
                        void* stream) {
  return fill_plane_with("sc_fill_plane", f, extents, dim, index, value, nullptr, stream);
}

sc_status sc_fill_plane_values(float* f, const int* extents, int dim, int index,
                               const float* values, void* stream) {
  if (values == nullptr) return fail(SC_ERR_INVALID, "sc_fill_plane_values: null values");
  return fill_plane_with("sc_fill_plane_values", f, extents, dim, index, 0.0f, values, stream);
}

sc_status sc_linear_interp(const float* v, const int* extents, int dim, float* out,
                           void* stream) {
  if (v == nullptr || extents == nullptr || out == nullptr)
    return fail(SC_ERR_INVALID, "sc_linear_interp: null argument");
  if (out == v) return fail(SC_ERR_INVALID, "sc_linear_interp: out must not alias v");
  if (dim < 0 || dim > 2)
    return fail(SC_ERR_INVALID, "sc_linear_interp: dim must be 0 (x), 1 (y) or 2 (z)");
  if (extents[0] < 1 || extents[1] < 1 || extents[2] < 1)
    return fail(SC_ERR_INVALID, "sc_linear_interp: extents must be positive");
  return hip_check(sc::launch_linear_interp(v, extents, dim, out, (hipStream_t)stream),
                   "linear interpolation kernel");
}

sc_status sc_riemann_flux(sc_ctx* c, int dim, const float* const* cons_neg,
                         const float* const* cons_pos, const float* const* prim_neg,
                         const float* const* prim_pos, const float* pdyn_neg,
                         const float* pdyn_pos, float* const* flux, void* stream) {
  if (c == nullptr || cons_neg == nullptr || cons_pos == nullptr || prim_neg == nullptr ||
      prim_pos == nullptr || flux == nullptr)
    return fail(SC_ERR_INVALID, "sc_riemann_flux: null argument");
  if (dim < 0 || dim > 2)
    return fail(SC_ERR_INVALID, "sc_riemann_flux: dim must be 0 (x), 1 (y) or 2 (z)");
  if ((pdyn_neg == nullptr) != (pdyn_pos == nullptr))
    return fail(SC_ERR_INVALID, "sc_riemann_flux: give p_dyn on both sides or neither");
  sc::RiemannArgs a{};
  const int nv = 7 + (c->nf - 5);
  for (int f = 0; f < c->nf; ++f) {
    if (cons_neg[f] == nullptr || cons_pos[f] == nullptr || flux[f] == nullptr)
      return fail(SC_ERR_INVALID, "sc_riemann_flux: null field pointer");
    a.cl[f] = cons_neg[f];
    a.cr[f] = cons_pos[f];
    a.out[f] = flux[f];
  }
  for (int v = 0; v < nv; ++v) {
    if (prim_neg[v] == nullptr || prim_pos[v] == nullptr)
      return fail(SC_ERR_INVALID, "sc_riemann_flux: null primitive pointer");
    a.wl[v] = prim_neg[v];
    a.wr[v] = prim_pos[v];
  }
  a.pdl = pdyn_neg;
  a.pdr = pdyn_pos;
  for (int d = 0; d < 3; ++d) a.N[d] = c->N[d];
  a.total = c->cells_full;
  a.has_pe = c->has_pe;
  a.pe_x = c->pe_x;
  a.pe_y = c->pe_y;
  a.pe_z = c->pe_z;
  const int flux_kind = c->params.numeric_flux == SC_HLLC ? sc::kHllc : sc::kHll;
  return hip_check(sc::launch_riemann_flux(a, dim, flux_kind, c->nf - 5, (hipStream_t)stream),
                   "riemann flux kernel");
}

namespace {
sc_status check_extents_spacing(const int* extents, const double* h, const char* what) {
  if (extents[0] < 1 || extents[1] < 1 || extents[2] < 1)
    return fail(SC_ERR_INVALID, std::string(what) + ": extents must be positive");
  if (!(h[0] > 0 && h[1] > 0 && h[2] > 0))
    return fail(SC_ERR_INVALID, std::string(what) + ": grid spacing must be positive");
  return SC_OK;
}
}  // namespace

sc_status sc_shear_stress(const float* u, const float* v, const float* w,
                          const float* const* rho_face, const int* extents, const double* h,
                          double nu, float* const* tau, void* stream) {
  if (u == nullptr || v == nullptr || w == nullptr || rho_face == nullptr ||
      extents == nullptr || h == nullptr || tau == nullptr)
    return fail(SC_ERR_INVALID, "sc_shear_stress: null argument");
  sc_status st = check_extents_spacing(extents, h, "sc_shear_stress");
  if (st != SC_OK) return st;
  sc::ShearArgs a{};
  a.u = u;
  a.v = v;
  a.w = w;
  for (int d = 0; d < 3; ++d) {
    if (rho_face[d] == nullptr) return fail(SC_ERR_INVALID, "sc_shear_stress: null face density");
    a.rho_f[d] = rho_face[d];
    a.N[d] = extents[d];
    a.h[d] = (float)h[d];
  }
  for (int t = 0; t < 9; ++t) {
    if (tau[t] == nullptr) return fail(SC_ERR_INVALID, "sc_shear_stress: null output");
    a.tau[t] = tau[t];
  }
  a.nu = (float)nu;
  return hip_check(sc::launch_shear_stress(a, (hipStream_t)stream), "shear stress kernel");
}

sc_status sc_heat_flux(const float* rho, const float* p, const int* extents, const double* h,
                       double nu, double pr, float* const* flux, void* stream) {
  if (rho == nullptr || p == nullptr || extents == nullptr || h == nullptr || flux == nullptr)
    return fail(SC_ERR_INVALID, "sc_heat_flux: null argument");
  sc_status st = check_extents_spacing(extents, h, "sc_heat_flux");
  if (st != SC_OK) return st;
  if (!(pr > 0)) return fail(SC_ERR_INVALID, "sc_heat_flux: pr must be positive");
  sc::HeatArgs a{};
  a.rho = rho;
  a.p = p;
  for (int d = 0; d < 3; ++d) {
    if (flux[d] == nullptr) return fail(SC_ERR_INVALID, "sc_heat_flux: null output");
    a.flux[d] = flux[d];
    a.N[d] = extents[d];
    a.h[d] = (float)h[d];
  }
  a.nu = (float)nu;
  a.inv_pr = (float)(1.0 / pr);
  return hip_check(sc::launch_heat_flux(a, (hipStream_t)stream), "heat flux kernel");
}

sc_status sc_exchange_halos(float* u, const int* extents, int halo, const sc_bc_face* bc,
                           void* stream) {
  if (u == nullptr || extents == nullptr || bc == nullptr)
    return fail(SC_ERR_INVALID, "sc_exchange_halos: null argument");
  if (halo < 1 || halo > SC_MAX_HALO)
    return fail(SC_ERR_INVALID, "sc_exchange_halos: halo width must be 1..4");
  if (extents[0] < 1 || extents[1] < 1 || extents[2] < 1)
    return fail(SC_ERR_INVALID, "sc_exchange_halos: interior extents must be positive");
  sc::HaloArgs ha{};
  ha.q[0] = u;
  ha.nf = 1;
  ha.h = halo;
  for (int d = 0; d < 3; ++d) {
    ha.n[d] = extents[d];
    ha.N[d] = extents[d] + 2 * halo;
  }
  ha.st[0] = ha.N[1];
  ha.st[1] = 1;
  ha.st[2] = (long long)ha.N[0] * ha.N[1];
  bool all_periodic = true;
  for (int d = 0; d < 3; ++d)
    for (int sd = 0; sd < 2; ++sd) {
      const sc_bc_face& b = bc[2 * d + sd];
      if (b.type < SC_BC_NONE || b.type > SC_BC_NEUMANN || b.type == SC_BC_REFLECTIVE)
        return fail(SC_ERR_UNSUPPORTED, "sc_exchange_halos: unsupported BC type");
      if ((b.type == SC_BC_NONE) != (bc[2 * d + 1 - sd].type == SC_BC_NONE))
        return fail(SC_ERR_INVALID, "SC_BC_NONE must be given on both faces of a dim");
      if (b.type == SC_BC_PERIODIC && extents[d] < halo)
        return fail(SC_ERR_UNSUPPORTED,
                    "PERIODIC BC needs at least halo_width interior cells");
      if (b.n_layer_values != 0 && b.n_layer_values != halo)
        return fail(SC_ERR_INVALID, "per-layer DIRICHLET values must have halo_width entries");
      sc::BcFaceDev& o = ha.bc[0][d][sd];
      o.type = b.type;
      o.value = b.value;
      o.n_layer_values = b.n_layer_values;
      for (int i = 0; i < SC_MAX_HALO; ++i) o.layer_values[i] = b.layer_values[i];
      o.layer_planes = b.layer_planes;
      if (b.type != SC_BC_PERIODIC) all_periodic = false;
    }
  hipStream_t st = (hipStream_t)stream;
  if (all_periodic && sc::halo_periodic_threads(ha) < (1LL << 31))
    return hip_check(sc::launch_halo_periodic(ha, st), "halo kernel");
  for (int d = 0; d < 3; ++d) {
    sc_status r = hip_check(sc::launch_halo_dim(ha, d, st), "halo kernel");
    if (r != SC_OK) return r;
  }
  return SC_OK;
}

sc_status sc_hllc_branch_codes(sc_ctx* c, const float* const* q, int dim, uint8_t* codes,
                               void* stream) {
  if (c == nullptr || q == nullptr || codes == nullptr)
    return fail(SC_ERR_INVALID, "sc_hllc_branch_codes: null argument");
  if (dim < 0 || dim > 2) return fail(SC_ERR_INVALID, "sc_hllc_branch_codes: dim must be 0..2");
  sc::SweepArgs a = c->sweep;
  for (int f = 0; f < c->nf; ++f) a.q.p[f] = q[f];
  hipError_t e;
  const int ns = c->nf - 5;
  if (c->params.interpolation == SC_WENO_5) {
    e = sc::launch_branch_codes<sc::kWeno5>(a, dim, ns, codes, (hipStream_t)stream);
  } else if (c->params.interpolation == SC_WENO_3) {
    e = sc::launch_branch_codes<sc::kWeno3>(a, dim, ns, codes, (hipStream_t)stream);
  } else {
    e = sc::launch_branch_codes<sc::kMuscl>(a, dim, ns, codes, (hipStream_t)stream);
  }
  return hip_check(e, "branch code kernel");
}

sc_status sc_ctx_set_reference_state(sc_ctx* c, const float* p_ref, const float* rho_ref) {
  if (c == nullptr) return fail(SC_ERR_INVALID, "sc_ctx_set_reference_state: null context");
  if ((p_ref == nullptr) != (rho_ref == nullptr))
    return fail(SC_ERR_INVALID, "sc_ctx_set_reference_state: give both p_ref and rho_ref, or neither");
  c->sweep.p_ref = p_ref;
  c->sweep.rho_ref = rho_ref;
  return SC_OK;
}

sc_status sc_ctx_get_tuning(const sc_ctx* c, sc_tuning* out) {
  if (c == nullptr || out == nullptr) return fail(SC_ERR_INVALID, "sc_ctx_get_tuning: null argument");
  from_tuning(c->tune, out);
  return SC_OK;
}

sc_status sc_ctx_set_tuning(sc_ctx* c, const sc_tuning* t) {
  if (c == nullptr) return fail(SC_ERR_INVALID, "sc_ctx_set_tuning: null context");
  sc::SweepTuning nt;
  const sc_status st = to_tuning(t, &nt);
  if (st == SC_OK) c->tune = nt;
  return st;
}

sc_status sc_ctx_enable_adjoint(sc_ctx* c) {
  if (c == nullptr) return fail(SC_ERR_INVALID, "sc_ctx_enable_adjoint: null context");
  if (c->nf > 5 && c->bbc_dims != 0)
    return fail(SC_ERR_UNSUPPORTED,
                "the adjoint takes boundary face / flux values without transported scalars only");
  if (c->adj_lq[0] != nullptr) return SC_OK;
  sc_status st = SC_OK;
  const size_t bytes = sizeof(float) * c->cells_full;
  for (int f = 0; f < c->nf && st == SC_OK; ++f)
    st = hip_check(hipMalloc((void**)&c->adj_lq[f], bytes), "hipMalloc adjoint");
  for (int f = 0; f < 9 && st == SC_OK; ++f)
    st = hip_check(hipMalloc((void**)&c->adj_lG[f], bytes), "hipMalloc adjoint");
  if (st != SC_OK) {
    for (int f = 0; f < SC_MAX_FIELDS; ++f) {
      if (c->adj_lq[f]) hipFree(c->adj_lq[f]);
      c->adj_lq[f] = nullptr;
    }
    for (int f = 0; f < 9; ++f) {
      if (c->adj_lG[f]) hipFree(c->adj_lG[f]);
      c->adj_lG[f] = nullptr;
    }
  }
  return st;
}

}  // extern "C"

namespace {

// lq = (dR/dq)^T g over the full grid (zero halo cells); q's halos are filled.
sc_status run_rhs_vjp(sc_ctx* c, float* const* q, const float* const* g, float* const* lq,
                      hipStream_t s) {
  sc_status st = sc_fill_halos(c, q, s);
  if (st != SC_OK) return st;
  const bool visc = c->params.include_diffusion != 0;
  float* zero[SC_MAX_FIELDS + 9];
  int nz = 0;
  for (int f = 0; f < c->nf; ++f) zero[nz++] = lq[f];
  if (visc)
    for (int f = 0; f < 9; ++f) zero[nz++] = c->adj_lG[f];
  st = hip_check(sc::launch_adj_zero(zero, nz, c->cells_full, s), "adjoint clear");
  if (st != SC_OK) return st;
  sc::AdjArgs A;
  memset(&A, 0, sizeof(A));
  A.s = c->sweep;
  for (int f = 0; f < c->nf; ++f) {
    A.s.q.p[f] = q[f];
    A.g.p[f] = g[f];
    A.lq.p[f] = lq[f];
  }
  for (int f = 0; f < 9; ++f) A.lG[f] = c->adj_lG[f];
  A.s.rk_form = 0;
  hipError_t e;
  const int scheme = c->params.interpolation;
  const int flux = c->params.numeric_flux;
  const int ns = c->nf - 5;
  if (scheme == SC_WENO_5) {
    e = sc::launch_adj_sweeps<sc::kWeno5>(A, c->h, flux, visc, ns, c->tune, s);
  } else if (scheme == SC_WENO_3) {
    e = sc::launch_adj_sweeps<sc::kWeno3>(A, c->h, flux, visc, ns, c->tune, s);
  } else {
    e = sc::launch_adj_sweeps<sc::kMuscl>(A, c->h, flux, visc, ns, c->tune, s);
  }
  st = hip_check(e, "adjoint sweep kernels");
  if (st != SC_OK) return st;
  if (visc) {
    st = hip_check(sc::launch_adj_visc_gather(A, s), "adjoint viscous gather");
    if (st != SC_OK) return st;
  }
  sc::HaloArgs ha = c->halo;
  for (int f = 0; f < c->nf; ++f) ha.q[f] = lq[f];
  if (ha.prim_mask != 0) {
    st = hip_check(sc::launch_halo_adj_prim(ha, q, s), "adjoint halo prim kernel");
    if (st != SC_OK) return st;
  }
  for (int d = 2; d >= 0; --d) {
    st = hip_check(sc::launch_halo_adj_dim(ha, q, d, s), "adjoint halo kernel");
    if (st != SC_OK) return st;
  }
  return SC_OK;
}

sc_status check_adjoint_ready(sc_ctx* c, const char* who) {
  if (c->adj_lq[0] == nullptr)
    return fail(SC_ERR_INVALID, std::string(who) + ": call sc_ctx_enable_adjoint first");
  if (c->nf > 5 && c->sweep.p_ref != nullptr)
    return fail(SC_ERR_UNSUPPORTED, std::string(who) +
                                        ": the adjoint takes a reference state without"
                                        " transported scalars only");
  return SC_OK;
}

}  // namespace

extern "C" {

sc_status sc_rhs_vjp(sc_ctx* c, float* const* q, const float* const* g, float* const* lq,
                     void* stream) {
  if (c == nullptr || q == nullptr || g == nullptr || lq == nullptr)
    return fail(SC_ERR_INVALID, "sc_rhs_vjp: null argument");
  sc_status st = check_adjoint_ready(c, "sc_rhs_vjp");
  if (st != SC_OK) return st;
  for (int f = 0; f < c->nf; ++f) {
    if (q[f] == nullptr || g[f] == nullptr || lq[f] == nullptr)
      return fail(SC_ERR_INVALID, "sc_rhs_vjp: null field pointer");
    for (int k = 0; k < c->nf; ++k)
      if (lq[f] == q[k] || lq[f] == g[k])
        return fail(SC_ERR_INVALID, "sc_rhs_vjp: lq must not alias q or g");
  }
  return run_rhs_vjp(c, q, g, lq, (hipStream_t)stream);
}

sc_status sc_rk_stage_vjp(sc_ctx* c, int form, float c0, float c1, const float* dt_dev,
                          float* const* uk, const float* const* lam_out, float* const* lam_uk,
                          float* const* lam_u0, void* stream) {
  if (c == nullptr || uk == nullptr || lam_out == nullptr || lam_uk == nullptr ||
      dt_dev == nullptr)
    return fail(SC_ERR_INVALID, "sc_rk_stage_vjp: null argument");
  sc_status st = check_adjoint_ready(c, "sc_rk_stage_vjp");
  if (st != SC_OK) return st;
  if (form != SC_RK_EULER && form != SC_RK_BLEND)
    return fail(SC_ERR_INVALID, "sc_rk_stage_vjp: form must be SC_RK_EULER or SC_RK_BLEND");
  for (int f = 0; f < c->nf; ++f) {
    if (uk[f] == nullptr || lam_out[f] == nullptr || lam_uk[f] == nullptr)
      return fail(SC_ERR_INVALID, "sc_rk_stage_vjp: null field pointer");
    if (lam_u0 != nullptr && lam_u0[f] == nullptr)
      return fail(SC_ERR_INVALID, "sc_rk_stage_vjp: null lam_u0 field pointer");
    for (int k = 0; k < c->nf; ++k) {
      if (lam_uk[f] == uk[k] || (lam_u0 != nullptr && (lam_u0[f] == uk[k] ||
                                                        lam_u0[f] == lam_out[k] ||
                                                        lam_u0[f] == lam_uk[k])))
        return fail(SC_ERR_INVALID, "sc_rk_stage_vjp: overlapping output arrays");
    }
  }
  hipStream_t s = (hipStream_t)stream;
  st = run_rhs_vjp(c, uk, lam_out, c->adj_lq, s);
  if (st != SC_OK) return st;
  sc::AdjGeom geo;
  geo.h = c->grid.halo;
  geo.n[0] = c->grid.nx;
  geo.n[1] = c->grid.ny;
  geo.n[2] = c->grid.nz;
  for (int d = 0; d < 3; ++d) geo.N[d] = c->N[d];
  const bool blend = form == SC_RK_BLEND;
  return hip_check(sc::launch_rk_adj_combine(lam_out, c->adj_lq, lam_uk,
                                             blend ? lam_u0 : nullptr, c->nf, geo,
                                             blend ? c0 : 0.0f, blend ? c1 : 1.0f, dt_dev, s),
                   "adjoint RK combine");
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched ensembles.
// ---------------------------------------------------------------------------

#include <vector>

#include "sc_batch_impl.hpp"

namespace sc {
template <int SCHEME>
hipError_t launch_batch_sweeps(const BatchArgs& b, const BatchTiles& tiles, int flux,
                               bool visc, int ns, bool full, const SweepTuning& tune,
                               hipStream_t stream);
}

struct sc_batch {
  sc_params params;
  int h = 0;
  int n_inst = 0;
  int nf = 5;
  long long total = 0;  // arena length (elements)
  sc::BatchInst* inst_dev = nullptr;
  sc::BatchHaloInst* hinst_dev = nullptr;
  sc::BatchTile* tiles_dev[3] = {nullptr, nullptr, nullptr};
  sc::BatchTiles tiles;
  sc::BatchHaloTile* htiles_dev[3] = {nullptr, nullptr, nullptr};
  int nhtiles[3] = {0, 0, 0};
  sc::BatchHaloTile* stiles_dev = nullptr;
  int nstiles = 0;
  float* hmin_dev = nullptr;
  unsigned* speed_bits = nullptr;
  float* r_arena[SC_MAX_FIELDS] = {};
  float* pe_dev = nullptr;                 // potential tables of every instance
  sc::BoundaryBc* bbc_dev = nullptr;       // boundary face / flux values
  sc::BatchHaloTile* ptiles_dev = nullptr; // primitive-BC finish tiles
  int nptiles = 0;
  bool full = false;                       // gravity, primitive BCs or boundary values
  // Every instance's halo fill is a pure copy along every dim
  // (copy_fill_dims): a stage's last sweep can write out's halos as images.
  bool images_ok = false;
  int img_clamp = 0;
  sc::HaloArgs halo_cfg;
  sc::SweepArgs proto;
  sc::SweepTuning tune;
  std::vector<sc_instance> inst_host;  // instance geometry, for rebuilding tiles
};

namespace {

template <typename T>
sc_status upload(const std::vector<T>& v, T** dst) {
  if (v.empty()) {
    *dst = nullptr;
    return SC_OK;
  }
  sc_status st = hip_check(hipMalloc((void**)dst, sizeof(T) * v.size()), "hipMalloc batch table");
  if (st != SC_OK) return st;
  return hip_check(hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice),
                   "hipMemcpy batch table");
}

// One block per tile: blocks pick up their tile descriptor in parallel (a
// grid-stride walk by fewer blocks serialises the descriptor loads, which
// dominated the halo fills' many small tiles).
int grid_for(int ntiles) { return ntiles > 0 ? ntiles : 1; }

// The dispatcher hands consecutive blocks to the 8 XCDs in turn, and each
// XCD has its own L2.  Reorders a tile list (one block per tile) so that
// XCD k runs the k-th contiguous range of the original order: neighbouring
// tiles of an instance, which share halo rows and transverse-neighbour
// planes, then hit the same L2.  The same bijection as the single-grid
// sweeps' block order, applied to the list on the host.
template <typename T>
void xcd_order(std::vector<T>& v) {
  const int n = (int)v.size();
  const int q = n / sc::kNumXcd, r = n % sc::kNumXcd;
  std::vector<T> o(v.size());
  for (int lin = 0; lin < n; ++lin) {
    const int xcd = lin % sc::kNumXcd, slot = lin / sc::kNumXcd;
    o[lin] = v[(xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + slot];
  }
  v.swap(o);
}

// Sweep tile lists of a batch for its current tuning: x sweep (y chunks of
// 64, z groups, x segments); y sweep (groups of by_y windows of the z
// plane's run of rows, runs of seg_y z planes); z sweep (y chunks, x groups,
// z segments); in the
// XCD order above when the tuning's xcd_remap is set.  Replaces any
// previous lists.
sc_status build_sweep_tiles(sc_batch* b) {
  const int rad = radius_of(b->params.interpolation);
  const int outw = 64 - 2 * rad - 2;
  const sc::SweepTuning& t = b->tune;
  std::vector<sc::BatchTile> tl[3];
  for (int i = 0; i < b->n_inst; ++i) {
    const sc_instance& g = b->inst_host[i];
    for (int bz = 0; bz < (g.nx + t.seg_x - 1) / t.seg_x; ++bz)
      for (int by = 0; by < (g.nz + t.by_x - 1) / t.by_x; ++by)
        for (int bx = 0; bx < (g.ny + 63) / 64; ++bx) tl[0].push_back({i, bx, by, bz});
    const int nwin = sc::y_windows(g.nx, g.ny, b->h, outw);
    for (int bz = 0; bz < (g.nz + t.seg_y - 1) / t.seg_y; ++bz)
      for (int bx = 0; bx < (nwin + t.by_y - 1) / t.by_y; ++bx) tl[1].push_back({i, bx, 0, bz});
    for (int bz = 0; bz < (g.nz + t.seg_z - 1) / t.seg_z; ++bz)
      for (int by = 0; by < (g.nx + t.by_z - 1) / t.by_z; ++by)
        for (int bx = 0; bx < (g.ny + 63) / 64; ++bx) tl[2].push_back({i, bx, by, bz});
  }
  for (int d = 0; d < 3; ++d) {
    if (t.xcd_remap) xcd_order(tl[d]);
    if (b->tiles_dev[d]) hipFree(b->tiles_dev[d]);
    b->tiles_dev[d] = nullptr;
    const sc_status st = upload(tl[d], &b->tiles_dev[d]);
    if (st != SC_OK) return st;
    b->tiles.t[d] = b->tiles_dev[d];
    b->tiles.n[d] = (int)tl[d].size();
    b->tiles.grid[d] = grid_for(b->tiles.n[d]);
  }
  return SC_OK;
}


}  // namespace

extern "C" {

sc_status sc_batch_get_tuning(const sc_batch* b, sc_tuning* out) {
  if (b == nullptr || out == nullptr)
    return fail(SC_ERR_INVALID, "sc_batch_get_tuning: null argument");
  from_tuning(b->tune, out);
  return SC_OK;
}

sc_status sc_batch_set_tuning(sc_batch* b, const sc_tuning* t) {
  if (b == nullptr) return fail(SC_ERR_INVALID, "sc_batch_set_tuning: null batch");
  sc::SweepTuning nt;
  sc_status st = to_tuning(t, &nt);
  if (st != SC_OK) return st;
  b->tune = nt;
  return build_sweep_tiles(b);
}

sc_status sc_batch_destroy(sc_batch* b) {
  if (b == nullptr) return SC_OK;
  hipFree(b->inst_dev);
  hipFree(b->hinst_dev);
  for (int d = 0; d < 3; ++d) {
    hipFree(b->tiles_dev[d]);
    hipFree(b->htiles_dev[d]);
  }
  hipFree(b->stiles_dev);
  hipFree(b->hmin_dev);
  hipFree(b->pe_dev);
  hipFree(b->bbc_dev);
  hipFree(b->ptiles_dev);
  hipFree(b->speed_bits);
  for (int f = 0; f < SC_MAX_FIELDS; ++f) hipFree(b->r_arena[f]);
  delete b;
  return SC_OK;
}

sc_status sc_batch_create(const sc_params* params, int halo, int n, const sc_instance* inst,
                          sc_batch** out) {
  if (params == nullptr || inst == nullptr || out == nullptr || n < 1)
    return fail(SC_ERR_INVALID, "sc_batch_create: null argument or no instances");
  *out = nullptr;
  const sc_params& p = *params;
  if (p.n_fields < 5 || p.n_fields > SC_MAX_FIELDS)
    return fail(SC_ERR_UNSUPPORTED, "n_fields must be 5, 6 or 7 (at most two transported scalars)");
  const int nf = p.n_fields;
  const int gdir = gravity_direction(p.g);
  if (p.gravity_source && gdir == -1)
    return fail(SC_ERR_INVALID, "Gravity direction must be specified if the gravity source"
                                " function is to be used.");
  for (int k = 0; k < p.n_boundary_values && k < SC_MAX_BOUNDARY_VALUES; ++k)
    if (p.boundary_values[k].plane != nullptr)
      return fail(SC_ERR_UNSUPPORTED,
                  "batched ensembles take scalar face / flux boundary values only (a plane"
                  " has the shape of one grid)");
  sc::BoundaryBc bbc[3][2];
  int bbc_dims = 0;
  {
    const sc_status bst = build_boundary_tables(p, bbc, &bbc_dims);
    if (bst != SC_OK) return bst;
  }
  // Validate the shared configuration with a throwaway single-grid check of
  // every instance shape (schemes, halo width, BC types).
  long long end = 0;
  for (int i = 0; i < n; ++i) {
    const sc_instance& g = inst[i];
    sc_grid grid{};
    grid.nx = g.nx;
    grid.ny = g.ny;
    grid.nz = g.nz;
    grid.halo = halo;
    grid.dx = g.dx;
    grid.dy = g.dy;
    grid.dz = g.dz;
    if (p.interpolation < SC_WENO_3 || p.interpolation > SC_MUSCL)
      return fail(SC_ERR_UNSUPPORTED, "interpolation_scheme is not implemented");
    if (p.numeric_flux != SC_HLL && p.numeric_flux != SC_HLLC)
      return fail(SC_ERR_INVALID, "numeric_flux_scheme is not implemented");
    if (g.nx < 1 || g.ny < 1 || g.nz < 1 || !(g.dx > 0 && g.dy > 0 && g.dz > 0))
      return fail(SC_ERR_INVALID, "instance extents and spacings must be positive");
    if (halo < radius_of(p.interpolation) + 1 || halo > SC_MAX_HALO)
      return fail(SC_ERR_INVALID, "halo_width too small for the interpolation stencil");
    if (g.nx >= (1 << 20) || (long long)(g.nx + 2 * halo) * (g.ny + 2 * halo) >= (1LL << 30))
      return fail(SC_ERR_UNSUPPORTED, "batched instance too large: nx must be below 2^20 and "
                                      "(nx + 2 * halo_width) * (ny + 2 * halo_width) below 2^30");
    if (gdir != -1 && (g.x_full == nullptr || g.y_full == nullptr || g.z_full == nullptr))
      return fail(SC_ERR_INVALID, "a gravity vector needs every instance's x_full, y_full and"
                                  " z_full coordinates");
    const int nn[3] = {g.nx, g.ny, g.nz};
    for (int f = 0; f < nf; ++f)
      for (int d = 0; d < 3; ++d)
        for (int s = 0; s < 2; ++s) {
          const sc_bc_face& bf = p.bc[f][d][s];
          if (bf.type < SC_BC_PERIODIC || bf.type > SC_BC_NEUMANN)
            return fail(SC_ERR_UNSUPPORTED, "unsupported BC type");
          if (bf.type == SC_BC_PERIODIC && nn[d] < halo)
            return fail(SC_ERR_UNSUPPORTED, "PERIODIC BC needs at least halo_width interior cells");
          if (bf.layer_planes != nullptr)
            return fail(SC_ERR_UNSUPPORTED, "batched ensembles do not support per-layer planes");
          if (bf.n_layer_values != 0 && bf.n_layer_values != halo)
            return fail(SC_ERR_INVALID, "per-layer DIRICHLET values must have halo_width entries");
        }
    const long long cells = (long long)(g.nx + 2 * halo) * (g.ny + 2 * halo) * (g.nz + 2 * halo);
    const long long NN[3] = {g.nx + 2LL * halo, g.ny + 2LL * halo, g.nz + 2LL * halo};
    for (int d = 0; d < 3; ++d)
      if (sc::halo_dim_threads(nf, halo, NN, d) > 0x7fffffffLL)
        return fail(SC_ERR_UNSUPPORTED, "batched instance too large (halo cells of a dim >= 2^31)");
    if (g.offset < end)
      return fail(SC_ERR_INVALID, "instance offsets must be increasing and non-overlapping");
    end = g.offset + cells;
  }

  sc_batch* b = new sc_batch();
  b->params = p;
  b->h = halo;
  b->nf = nf;
  b->n_inst = n;
  b->full = gdir != -1 || p.prim_bc_mask != 0 || bbc_dims != 0;
  b->total = end;
  b->inst_host.assign(inst, inst + n);

  std::vector<sc::BatchInst> bi(n);
  std::vector<sc::BatchHaloInst> hi(n);
  std::vector<float> hmin(n);
  std::vector<sc::BatchHaloTile> ht[3], stl, ptl;
  // Potential tables (gravity): per instance pe_x over NX, pe_y over NY,
  // pe_z over NZ, tabulated as in sc_ctx_create (generic.py:312-319).
  std::vector<float> pe_host;
  for (int i = 0; i < n; ++i) {
    const sc_instance& g = inst[i];
    const double hh[3] = {g.dx, g.dy, g.dz};
    sc::BatchInst& e = bi[i];
    e.nx = g.nx;
    e.ny = g.ny;
    e.nz = g.nz;
    e.off = g.offset;
    e.pe_off = -1;
    if (gdir != -1) {
      e.pe_off = (long long)pe_host.size();
      const float* coord[3] = {g.x_full, g.y_full, g.z_full};
      const int NN3[3] = {g.nx + 2 * halo, g.ny + 2 * halo, g.nz + 2 * halo};
      for (int d = 0; d < 3; ++d) {
        const double gv = 9.81 * p.g[d];
        const float coef = d == 0 ? (float)(-gv) : (float)gv;
        for (int c = 0; c < NN3[d]; ++c) pe_host.push_back(coef * coord[d][c]);
      }
    }
    for (int d = 0; d < 3; ++d) {
      e.h[d] = (float)hh[d];
      e.inv_h[d] = 1.0f / (float)hh[d];
      e.inv_two_h[d] = 1.0f / (float)(2.0 * hh[d]);
    }
    hi[i].n[0] = g.nx;
    hi[i].n[1] = g.ny;
    hi[i].n[2] = g.nz;
    hi[i].off = g.offset;
    hi[i].pe_off = e.pe_off;
    float hm = e.h[0];
    if (e.h[1] < hm) hm = e.h[1];
    if (e.h[2] < hm) hm = e.h[2];
    hmin[i] = hm;
    const long long N[3] = {g.nx + 2LL * halo, g.ny + 2LL * halo, g.nz + 2LL * halo};
    for (int d = 0; d < 3; ++d) {
      const long long threads = sc::halo_dim_threads(nf, halo, N, d);
      for (long long k = 0; k < (threads + 255) / 256; ++k) ht[d].push_back({i, (int)k});
    }
    if (p.prim_bc_mask != 0) {
      const long long padded = N[0] * N[1] * N[2];
      for (long long k = 0; k < (padded + 255) / 256; ++k) ptl.push_back({i, (int)k});
    }
    const long long cells = (long long)g.nx * g.ny * g.nz;
    for (long long k = 0; k < (cells + sc::kSpeedChunk - 1) / sc::kSpeedChunk; ++k)
      stl.push_back({i, (int)k});
  }
  sc_status st = upload(bi, &b->inst_dev);
  if (st == SC_OK) st = upload(hi, &b->hinst_dev);
  if (st == SC_OK) st = upload(hmin, &b->hmin_dev);
  if (st == SC_OK) st = build_sweep_tiles(b);
  for (int d = 0; d < 3 && st == SC_OK; ++d) {
    st = upload(ht[d], &b->htiles_dev[d]);
    b->nhtiles[d] = (int)ht[d].size();
  }
  if (st == SC_OK) st = upload(stl, &b->stiles_dev);
  b->nstiles = (int)stl.size();
  if (st == SC_OK) st = upload(ptl, &b->ptiles_dev);
  b->nptiles = (int)ptl.size();
  if (st == SC_OK) st = upload(pe_host, &b->pe_dev);
  if (st == SC_OK && bbc_dims != 0) {
    st = hip_check(hipMalloc((void**)&b->bbc_dev, sizeof(bbc)), "hipMalloc boundary values");
    if (st == SC_OK)
      st = hip_check(hipMemcpy(b->bbc_dev, bbc, sizeof(bbc), hipMemcpyHostToDevice),
                     "hipMemcpy boundary values");
  }
  if (st == SC_OK)
    st = hip_check(hipMalloc((void**)&b->speed_bits, sizeof(unsigned) * n), "hipMalloc speed");
  for (int f = 0; f < nf && st == SC_OK; ++f) {
    st = hip_check(hipMalloc((void**)&b->r_arena[f], sizeof(float) * b->total), "hipMalloc R");
    if (st == SC_OK)
      st = hip_check(hipMemset(b->r_arena[f], 0, sizeof(float) * b->total), "hipMemset R");
  }
  if (st != SC_OK) {
    sc_batch_destroy(b);
    return st;
  }
  // Shared halo configuration (BC tables).
  sc::HaloArgs& ha = b->halo_cfg;
  memset(&ha, 0, sizeof(ha));
  ha.nf = nf;
  ha.h = halo;
  ha.prim_mask = p.prim_bc_mask;
  for (int f = 0; f < nf; ++f)
    for (int d = 0; d < 3; ++d)
      for (int s = 0; s < 2; ++s) {
        const sc_bc_face& bf = p.bc[f][d][s];
        sc::BcFaceDev& o = ha.bc[f][d][s];
        o.type = bf.type;
        o.value = bf.value;
        o.n_layer_values = bf.n_layer_values;
        for (int k = 0; k < SC_MAX_HALO; ++k) o.layer_values[k] = bf.layer_values[k];
        o.layer_planes = nullptr;
      }
  // Shared sweep configuration.
  sc::SweepArgs& a = b->proto;
  memset(&a, 0, sizeof(a));
  a.h = halo;
  a.has_pe = 0;  // per instance (tile_args)
  a.grav_src = p.gravity_source ? 1 : 0;
  a.grav_mask = 0;
  if (p.gravity_source) a.grav_mask = gdir == 3 ? 7 : (1 << gdir);
  for (int d = 0; d < 3; ++d) a.gdir[d] = (float)p.g[d];
  a.prim_bc_mask = p.prim_bc_mask;
  a.bbc = b->bbc_dev;
  a.bbc_dims = bbc_dims;
  a.vc.nu = (float)p.nu;
  a.vc.kfac = (float)(1.0 / p.pr) * sc::kCp;
  // Halo images from the last sweep: the copy-fill condition must hold for
  // every instance's extents (its periodic dims at least 2h long).
  b->images_ok = true;
  for (int i = 0; i < n && b->images_ok; ++i) {
    const int n3[3] = {inst[i].nx, inst[i].ny, inst[i].nz};
    int clamp = 0;
    b->images_ok = copy_fill_dims(p, nf, n3, halo, &clamp);
    b->img_clamp = clamp;
  }
  *out = b;
  return SC_OK;
}

sc_status sc_batch_fill_halos(sc_batch* b, float* const* q, void* stream) {
  if (b == nullptr || q == nullptr) return fail(SC_ERR_INVALID, "sc_batch_fill_halos: null argument");
  for (int f = 0; f < b->nf; ++f)
    if (q[f] == nullptr) return fail(SC_ERR_INVALID, "sc_batch_fill_halos: null arena");
  for (int d = 0; d < 3; ++d) {
    sc_status st = hip_check(
        sc::launch_batch_halo_dim(b->halo_cfg, q, b->hinst_dev, b->htiles_dev[d], b->nhtiles[d],
                                  grid_for(b->nhtiles[d]), d, (hipStream_t)stream),
        "batch halo kernel");
    if (st != SC_OK) return st;
  }
  if (b->halo_cfg.prim_mask != 0)
    return hip_check(sc::launch_batch_halo_prim_finish(b->halo_cfg, q, b->hinst_dev,
                                                       b->ptiles_dev, b->nptiles,
                                                       grid_for(b->nptiles), (hipStream_t)stream),
                     "batch halo prim kernel");
  return SC_OK;
}

sc_status sc_batch_rk_stage_ex(sc_batch* b, int form, float c0, float c1,
                               const float* dt_dev, const float* const* u0, float* const* uk,
                               float* const* out, int flags, float* speed_dev, void* stream) {
  if (b == nullptr || uk == nullptr || out == nullptr || dt_dev == nullptr)
    return fail(SC_ERR_INVALID, "sc_batch_rk_stage_ex: null argument");
  if (form != SC_RK_EULER && form != SC_RK_BLEND)
    return fail(SC_ERR_UNSUPPORTED, "sc_batch_rk_stage_ex: form must be SC_RK_EULER or SC_RK_BLEND");
  if (form == SC_RK_BLEND && u0 == nullptr)
    return fail(SC_ERR_INVALID, "sc_batch_rk_stage_ex: u0 required");
  for (int f = 0; f < b->nf; ++f) {
    if (uk[f] == nullptr || out[f] == nullptr || (form == SC_RK_BLEND && u0[f] == nullptr))
      return fail(SC_ERR_INVALID, "sc_batch_rk_stage_ex: null arena");
    if (out[f] == uk[f]) return fail(SC_ERR_INVALID, "sc_batch_rk_stage_ex: out must not alias uk");
  }
  if (flags & ~(SC_STAGE_HALOS_READY | SC_STAGE_FILL_OUT))
    return fail(SC_ERR_INVALID, "sc_batch_rk_stage_ex: unknown flags");
  if (!(flags & SC_STAGE_HALOS_READY)) {
    const sc_status st = sc_batch_fill_halos(b, uk, stream);
    if (st != SC_OK) return st;
  }
  // out's halos: written by the last sweep as images where that is exactly
  // the fill (copy_fill_dims for every instance), else filled after.
  const bool fill_out = (flags & SC_STAGE_FILL_OUT) != 0;
  sc::BatchArgs ba;
  memset(&ba, 0, sizeof(ba));
  ba.proto = b->proto;
  ba.proto.rk_form = form;
  ba.proto.c0 = c0;
  ba.proto.c1 = c1;
  ba.proto.out_images = fill_out && b->images_ok;
  ba.proto.img_clamp = b->img_clamp;
  ba.nf = b->nf;
  ba.pe = b->pe_dev;
  for (int f = 0; f < b->nf; ++f) {
    ba.q.p[f] = uk[f];
    ba.u0.p[f] = form == SC_RK_BLEND ? u0[f] : nullptr;
    ba.r.p[f] = b->r_arena[f];
    ba.out.p[f] = out[f];
  }
  ba.inst = b->inst_dev;
  ba.dt = dt_dev;
  ba.speed = (unsigned*)speed_dev;
  hipError_t e;
  const int flux = b->params.numeric_flux;
  const bool visc = b->params.include_diffusion != 0;
  hipStream_t s = (hipStream_t)stream;
  const int ns = b->nf - 5;
  if (b->params.interpolation == SC_WENO_5) {
    e = sc::launch_batch_sweeps<sc::kWeno5>(ba, b->tiles, flux, visc, ns, b->full, b->tune, s);
  } else if (b->params.interpolation == SC_WENO_3) {
    e = sc::launch_batch_sweeps<sc::kWeno3>(ba, b->tiles, flux, visc, ns, b->full, b->tune, s);
  } else {
    e = sc::launch_batch_sweeps<sc::kMuscl>(ba, b->tiles, flux, visc, ns, b->full, b->tune, s);
  }
  const sc_status st = hip_check(e, "batch sweep kernels");
  if (st != SC_OK || !fill_out || ba.proto.out_images) return st;
  return sc_batch_fill_halos(b, out, stream);
}

sc_status sc_batch_rk_stage(sc_batch* b, int form, float c0, float c1, const float* dt_dev,
                            const float* const* u0, float* const* uk, float* const* out,
                            void* stream) {
  return sc_batch_rk_stage_ex(b, form, c0, c1, dt_dev, u0, uk, out, 0, nullptr, stream);
}

sc_status sc_batch_cfl_dt(sc_batch* b, float* speed_dev, float cfl, float* dt_dev,
                          void* stream) {
  if (b == nullptr || speed_dev == nullptr || dt_dev == nullptr)
    return fail(SC_ERR_INVALID, "sc_batch_cfl_dt: null argument");
  return hip_check(sc::launch_batch_cfl_dt_reset((unsigned*)speed_dev, b->n_inst, cfl,
                                                 b->hmin_dev, dt_dev, (hipStream_t)stream),
                   "batch cfl dt kernel");
}

sc_status sc_batch_max_wave_speed(sc_batch* b, const float* const* q, float cfl,
                                  float* speed_dev, float* dt_dev, void* stream) {
  if (b == nullptr || q == nullptr || speed_dev == nullptr)
    return fail(SC_ERR_INVALID, "sc_batch_max_wave_speed: null argument");
  float* qb[SC_MAX_FIELDS] = {};
  for (int f = 0; f < 5; ++f) qb[f] = const_cast<float*>(q[f]);
  return hip_check(sc::launch_batch_max_speed(qb, b->hinst_dev, b->h, b->stiles_dev, b->nstiles,
                                              grid_for(b->nstiles), b->n_inst,
                                              (unsigned*)speed_dev, cfl, b->hmin_dev, dt_dev,
                                              b->pe_dev, (hipStream_t)stream),
                   "batch max speed kernel");
}

}  // extern "C"
