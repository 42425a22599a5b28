// capi.cpp -- the extern "C" entry points of liblbm_mi355x.so (declared in include/lbm_mi355x.h).
#include <new>
#include <string>

#include "handle.h"

namespace {
thread_local std::string g_create_error;

lbm::SpSolver *as_sp(lbm_handle *h) { return (h && h->kind == lbm_handle::SP) ? static_cast<lbm::SpSolver *>(h) : nullptr; }
const lbm::SpSolver *as_sp(const lbm_handle *h)
{
    return (h && h->kind == lbm_handle::SP) ? static_cast<const lbm::SpSolver *>(h) : nullptr;
}
lbm::CgSolver *as_cg(lbm_handle *h) { return (h && h->kind == lbm_handle::CG) ? static_cast<lbm::CgSolver *>(h) : nullptr; }
const lbm::CgSolver *as_cg(const lbm_handle *h)
{
    return (h && h->kind == lbm_handle::CG) ? static_cast<const lbm::CgSolver *>(h) : nullptr;
}
lbm::StSolver *as_st(lbm_handle *h) { return (h && h->kind == lbm_handle::ST) ? static_cast<lbm::StSolver *>(h) : nullptr; }
const lbm::StSolver *as_st(const lbm_handle *h)
{
    return (h && h->kind == lbm_handle::ST) ? static_cast<const lbm::StSolver *>(h) : nullptr;
}
}  // namespace

void lbm::set_create_error(const std::string &msg) { g_create_error = msg; }

extern "C" {

const char *lbm_version(void) { return "lbm-mi355x 0.1 (gfx950, sparse D3Q19 MRT)"; }

const char *lbm_last_error(const lbm_handle *h)
{
    if (!h) return g_create_error.c_str();
    // a copy per calling thread: the handle's message may be replaced by another thread's call
    thread_local std::string msg;
    msg = h->error();
    return msg.c_str();
}

void lbm_destroy(lbm_handle *h) { delete h; }

int lbm_geometry_wraps(int32_t nx, int32_t ny, int32_t nz, const int32_t *solid_ijk, int32_t axis, int32_t *wraps)
{
    if (!wraps || axis < 0 || axis > 2) {
        g_create_error = "lbm_geometry_wraps: axis must be 0, 1 or 2 and wraps non-null";
        return LBM_ERR_ARG;
    }
    lbm::SparseGeometry g;
    std::string err;
    const int rc = lbm::build_sparse_geometry(nx, ny, nz, solid_ijk, g, err);
    if (rc != LBM_OK) {
        g_create_error = "lbm_geometry_wraps: " + err;
        return rc;
    }
    *wraps = lbm::pores_wrap_along(g, axis) ? 1 : 0;
    return LBM_OK;
}

int lbm_sp_create(const lbm_sp_params *p, const int32_t *solid_ijk, lbm_handle **out)
{
    if (!p || !solid_ijk || !out) {
        g_create_error = "lbm_sp_create: null argument";
        return LBM_ERR_ARG;
    }
    *out = nullptr;
    auto *s = new (std::nothrow) lbm::SpSolver();
    if (!s) {
        g_create_error = "lbm_sp_create: out of host memory";
        return LBM_ERR_NOMEM;
    }
    int rc = s->create(*p, solid_ijk);
    if (rc != LBM_OK) {
        g_create_error = s->error();
        delete s;
        return rc;
    }
    *out = s;
    return LBM_OK;
}

int lbm_sp_count(const lbm_handle *h, int64_t *count)
{
    auto *s = as_sp(h);
    if (!s || !count) return LBM_ERR_ARG;
    *count = s->count();
    return LBM_OK;
}

int lbm_sp_set_boundaries(lbm_handle *h, const lbm_sp_bc *bc)
{
    auto *s = as_sp(h);
    return s ? s->set_boundaries(bc) : LBM_ERR_ARG;
}

int lbm_sp_set_power_law(lbm_handle *h, double m, double n)
{
    auto *s = as_sp(h);
    return s ? s->set_power_law(m, n) : LBM_ERR_ARG;
}

int lbm_sp_get_viscosity(lbm_handle *h, double *nu)
{
    auto *s = as_sp(h);
    return s ? s->get_viscosity(nu) : LBM_ERR_ARG;
}

int lbm_sp_get_dense(lbm_handle *h, double *u_ijk, double *rho_ijk)
{
    auto *s = as_sp(h);
    return s ? s->get_dense(u_ijk, rho_ijk) : LBM_ERR_ARG;
}

int lbm_sp_init(lbm_handle *h, double rho, const double u0[3])
{
    auto *s = as_sp(h);
    return s ? s->init(rho, u0) : LBM_ERR_ARG;
}

int lbm_sp_load_state(lbm_handle *h, const double *f, const double *u, const double *rho)
{
    auto *s = as_sp(h);
    return s ? s->load_state(f, u, rho) : LBM_ERR_ARG;
}

int lbm_sp_step(lbm_handle *h, int32_t nsteps)
{
    auto *s = as_sp(h);
    return s ? s->step(nsteps, nullptr) : LBM_ERR_ARG;
}

int lbm_sp_step_timed(lbm_handle *h, int32_t nsteps, double *elapsed_ms)
{
    auto *s = as_sp(h);
    if (!s || !elapsed_ms) return LBM_ERR_ARG;
    return s->step(nsteps, elapsed_ms);
}

int lbm_sp_get_state(lbm_handle *h, double *f, double *u, double *rho)
{
    auto *s = as_sp(h);
    return s ? s->get_state(f, u, rho) : LBM_ERR_ARG;
}

int lbm_sp_get_state_noalloc(lbm_handle *h, double *f, double *u, double *rho)
{
    auto *s = as_sp(h);
    return s ? s->get_state_noalloc(f, u, rho) : LBM_ERR_ARG;
}

int lbm_sp_error_ranks(lbm_handle *h, int32_t nranks, double out[3])
{
    auto *s = as_sp(h);
    return s ? s->error_ranks(nranks, out) : LBM_ERR_ARG;
}

int lbm_sp_cube_sums(lbm_handle *h, const int32_t origin[3], int32_t size, int32_t m, double *out)
{
    auto *s = as_sp(h);
    return s ? s->cube_sums(origin, size, m, out) : LBM_ERR_ARG;
}

int lbm_get_launch(const lbm_handle *h, int32_t kernel, lbm_launch *out)
{
    if (auto *s = as_sp(h)) return s->get_launch(kernel, out);
    if (auto *c = as_cg(h)) return c->get_launch(kernel, out);
    if (auto *t = as_st(h)) return t->get_launch(kernel, out);
    return LBM_ERR_ARG;
}

int lbm_set_launch(lbm_handle *h, int32_t kernel, const lbm_launch *in)
{
    if (auto *s = as_sp(h)) return s->set_launch(kernel, in);
    if (auto *c = as_cg(h)) return c->set_launch(kernel, in);
    if (auto *t = as_st(h)) return t->set_launch(kernel, in);
    return LBM_ERR_ARG;
}

int lbm_sp_diagnostics(lbm_handle *h, const int32_t *box, lbm_sp_diag *out)
{
    auto *s = as_sp(h);
    return s ? s->diagnostics(box, out) : LBM_ERR_ARG;
}

int lbm_sp_steps_done(const lbm_handle *h, int64_t *steps)
{
    auto *s = as_sp(h);
    if (!s || !steps) return LBM_ERR_ARG;
    *steps = s->steps_done();
    return LBM_OK;
}

int lbm_cg_create(const lbm_cg_params *p, const int32_t *solid_ijk, lbm_handle **out)
{
    if (!p || !solid_ijk || !out) {
        g_create_error = "lbm_cg_create: null argument";
        return LBM_ERR_ARG;
    }
    *out = nullptr;
    auto *s = new (std::nothrow) lbm::CgSolver();
    if (!s) {
        g_create_error = "lbm_cg_create: out of host memory";
        return LBM_ERR_NOMEM;
    }
    int rc = s->create(*p, solid_ijk);
    if (rc != LBM_OK) {
        g_create_error = s->error();
        delete s;
        return rc;
    }
    *out = s;
    return LBM_OK;
}

int lbm_cg_count(const lbm_handle *h, int64_t *count)
{
    auto *s = as_cg(h);
    if (!s || !count) return LBM_ERR_ARG;
    *count = s->count();
    return LBM_OK;
}

int lbm_cg_init(lbm_handle *h, const double *psi_ijk, const double u0[3])
{
    auto *s = as_cg(h);
    return s ? s->init(psi_ijk, u0) : LBM_ERR_ARG;
}

int lbm_cg_load_state(lbm_handle *h, const double *f, const double *u, const double *rho, const double *psi,
                      const double *rho_r, const double *rho_b)
{
    auto *s = as_cg(h);
    return s ? s->load_state(f, u, rho, psi, rho_r, rho_b) : LBM_ERR_ARG;
}

int lbm_cg_step(lbm_handle *h, int32_t nsteps)
{
    auto *s = as_cg(h);
    return s ? s->step(nsteps, nullptr) : LBM_ERR_ARG;
}

int lbm_cg_step_timed(lbm_handle *h, int32_t nsteps, double *elapsed_ms)
{
    auto *s = as_cg(h);
    if (!s || !elapsed_ms) return LBM_ERR_ARG;
    return s->step(nsteps, elapsed_ms);
}

int lbm_cg_set_body_force(lbm_handle *h, const double g[3])
{
    auto *s = as_cg(h);
    return s ? s->set_body_force(g) : LBM_ERR_ARG;
}

int lbm_cg_get_state(lbm_handle *h, double *f, double *u, double *rho, double *psi, double *rho_r, double *rho_b)
{
    auto *s = as_cg(h);
    return s ? s->get_state(f, u, rho, psi, rho_r, rho_b) : LBM_ERR_ARG;
}

int lbm_cg_diagnostics(lbm_handle *h, const int32_t *box, lbm_cg_diag *out)
{
    auto *s = as_cg(h);
    return s ? s->diagnostics(box, out) : LBM_ERR_ARG;
}

int lbm_cg_steps_done(const lbm_handle *h, int64_t *steps)
{
    auto *s = as_cg(h);
    if (!s || !steps) return LBM_ERR_ARG;
    *steps = s->steps_done();
    return LBM_OK;
}

int lbm_st_create(const lbm_st_params *p, const int32_t *solid_ijk, lbm_handle **out)
{
    if (!p || !solid_ijk || !out) {
        g_create_error = "lbm_st_create: null argument";
        return LBM_ERR_ARG;
    }
    *out = nullptr;
    auto *s = new (std::nothrow) lbm::StSolver();
    if (!s) {
        g_create_error = "lbm_st_create: out of host memory";
        return LBM_ERR_NOMEM;
    }
    int rc = s->create(*p, solid_ijk);
    if (rc != LBM_OK) {
        g_create_error = s->error();
        delete s;
        return rc;
    }
    *out = s;
    return LBM_OK;
}

int lbm_st_create_coupled(const lbm_st_params *p, lbm_handle *flow, lbm_handle **out)
{
    auto *f = as_sp(flow);
    if (!p || !f || !out) {
        g_create_error = "lbm_st_create_coupled: null argument or flow handle is not single-phase";
        return LBM_ERR_ARG;
    }
    *out = nullptr;
    auto *s = new (std::nothrow) lbm::StSolver();
    if (!s) {
        g_create_error = "lbm_st_create_coupled: out of host memory";
        return LBM_ERR_NOMEM;
    }
    int rc = s->create_coupled(*p, f);
    if (rc != LBM_OK) {
        g_create_error = s->error();
        delete s;
        return rc;
    }
    *out = s;
    return LBM_OK;
}

int lbm_st_count(const lbm_handle *h, int64_t *count)
{
    auto *s = as_st(h);
    if (!s || !count) return LBM_ERR_ARG;
    *count = s->count();
    return LBM_OK;
}

int lbm_st_set_velocity(lbm_handle *h, const double *u)
{
    auto *s = as_st(h);
    return s ? s->set_velocity(u) : LBM_ERR_ARG;
}

int lbm_st_init(lbm_handle *h, const double *conc_ijk)
{
    auto *s = as_st(h);
    return s ? s->init(conc_ijk) : LBM_ERR_ARG;
}

int lbm_st_load_state(lbm_handle *h, const double *g)
{
    auto *s = as_st(h);
    return s ? s->load_state(g) : LBM_ERR_ARG;
}

int lbm_st_step(lbm_handle *h, int32_t nsteps)
{
    auto *s = as_st(h);
    return s ? s->step(nsteps, nullptr) : LBM_ERR_ARG;
}

int lbm_st_step_timed(lbm_handle *h, int32_t nsteps, double *elapsed_ms)
{
    auto *s = as_st(h);
    if (!s || !elapsed_ms) return LBM_ERR_ARG;
    return s->step(nsteps, elapsed_ms);
}

int lbm_st_get_state(lbm_handle *h, double *g, double *conc)
{
    auto *s = as_st(h);
    return s ? s->get_state(g, conc) : LBM_ERR_ARG;
}

int lbm_st_diagnostics(lbm_handle *h, lbm_st_diag *out)
{
    auto *s = as_st(h);
    return s ? s->diagnostics(out) : LBM_ERR_ARG;
}

}  // extern "C"
