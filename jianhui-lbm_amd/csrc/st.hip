// st.hip -- passive solute: D3Q7 BGK advection-diffusion on the sparse fluid nodes.
//
// The reference's six-velocity solute program (Single_Phase_3D/
// Solute_Transport_MPI_Software_Spars_6V.cpp, ST6) relaxes g_q = C/6 (1 + 3 e_q.u) with one rate
// 1/tau_s, tau_s = D/c_s^2 + 1/2 (ST6:1274, 1466-1474), pushes to the downstream node or bounces
// back into the node's opposite population at a solid site (ST6:1481-1535, zero flux), and
// sums the populations into C (ST6:1625-1640).  Here the lattice carries a rest population of
// weight w0: w_q = (1 - w0)/6 on the six axis directions, c_s^2 = (1 - w0)/3,
// g_eq_q = w_q C (1 + e_q.u / c_s^2); w0 = 0 is the six-velocity scheme itself.
//
// One launch per step: every thread owns one node, pulls its 7 populations (the axis rows of
// the packed D3Q19 pull table are the D3Q7 links), relaxes and writes the post-collision
// populations.  As in the flow kernels the stored state is post-collision; get_state pulls.
//
// Data layout (row stride P): pops[2] 7 x P, u 3 x P (the flow handle's when coupled),
// c_init P (the concentration of an equilibrium start).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"
#include "handle.h"
#include "lattice.h"
#include "nbrpack.h"
#include "pull.h"
#include "storage.h"
#include "tune.h"

namespace lbm {

constexpr int Q7 = 7;

template <typename C>
struct StArgs {
    const C *__restrict__ src;
    C *__restrict__ dst;
    NbrDev nbr;
    const C *__restrict__ u;
    const C *__restrict__ cinit;
    int64_t N, P;
    C omega;     // 1/tau_s
    C w0, wq;    // rest and axis weights
    C inv_cs2;
};

// e_q.u for the axis direction q = 1..6 (+x, -x, +y, -y, +z, -z: the first six D3Q19 rows)
template <typename C>
__device__ __forceinline__ C axis_dot(int q, C ux, C uy, C uz)
{
    return q == 1 ? ux : q == 2 ? -ux : q == 3 ? uy : q == 4 ? -uy : q == 5 ? uz : -uz;
}

// the 7 post-streaming populations of node n
template <typename C, int MODE>
__device__ __forceinline__ void st_gather(const StArgs<C> &a, int64_t n, C (&g)[Q7])
{
    g[0] = a.src[n];
    if (MODE == K_DIRECT) {
#pragma unroll
        for (int q = 1; q < Q7; ++q) g[q] = a.src[(int64_t)q * a.P + n];
        return;
    }
    uint32_t nb[6];
    load_nbr_first<6>(a.nbr, a.P, n, nb);
#pragma unroll
    for (int q = 1; q < Q7; ++q) {
        const uint32_t v = nb[q - 1];
        const int64_t row = (v & BB_FLAG) ? (int64_t)opp(q) : (int64_t)q;
        g[q] = a.src[row * a.P + (int64_t)(v & ~BB_FLAG)];
    }
}

template <typename C, int MODE, int NB = BLOCK>
__global__ void __launch_bounds__(NB) st_step_kernel(StArgs<C> a, int32_t order)
{
    const int64_t n = tile_of(order, blockIdx.x, gridDim.x) * NB + threadIdx.x;
    if (n >= a.N) return;
    const C ux = a.u[n], uy = a.u[a.P + n], uz = a.u[2 * a.P + n];
    C g[Q7], c;
    if (MODE == K_INIT) {
        c = a.cinit[n];   // g = g_eq(C, u): its collision leaves it unchanged
    } else {
        st_gather<C, MODE>(a, n, g);
        c = g[0];
#pragma unroll
        for (int q = 1; q < Q7; ++q) c += g[q];
    }
    const C wc = a.wq * c;
    {
        const C eq = a.w0 * c;
        const C v = MODE == K_INIT ? eq : g[0] - a.omega * (g[0] - eq);
        __builtin_nontemporal_store(v, &a.dst[n]);
    }
#pragma unroll
    for (int q = 1; q < Q7; ++q) {
        const C eq = wc * (C(1) + axis_dot(q, ux, uy, uz) * a.inv_cs2);
        const C v = MODE == K_INIT ? eq : g[q] - a.omega * (g[q] - eq);
        __builtin_nontemporal_store(v, &a.dst[(int64_t)q * a.P + n]);
    }
}

// post-streaming populations and C in the reference's layout (g[count][7], C[count])
// of the nodes n0 <= n < n1, at local index n - n0
template <typename C, int MODE>
__global__ void __launch_bounds__(BLOCK) st_get_state_kernel(StArgs<C> a, int64_t n0, int64_t n1,
                                                             double *__restrict__ g_out, double *__restrict__ c_out)
{
    const int64_t n = group_floor(n0) + (int64_t)blockIdx.x * BLOCK + threadIdx.x;   // see load_nbr
    if (n < n0 || n >= n1) return;
    const int64_t l = n - n0;
    C g[Q7];
    if (MODE == K_INIT) {
        const C c = a.cinit[n], wc = a.wq * c;
        const C ux = a.u[n], uy = a.u[a.P + n], uz = a.u[2 * a.P + n];
        g[0] = a.w0 * c;
#pragma unroll
        for (int q = 1; q < Q7; ++q) g[q] = wc * (C(1) + axis_dot(q, ux, uy, uz) * a.inv_cs2);
    } else {
        st_gather<C, MODE>(a, n, g);
    }
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < Q7; ++q) {
        if (g_out) g_out[l * Q7 + q] = (double)g[q];
        s += (double)g[q];
    }
    if (c_out) c_out[l] = s;
}

template <typename C>
__global__ void __launch_bounds__(BLOCK) st_load_state_kernel(StArgs<C> a, int64_t n0, int64_t n1,
                                                              const double *__restrict__ g_in)
{
    const int64_t n = n0 + (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (n >= n1) return;
#pragma unroll
    for (int q = 0; q < Q7; ++q) a.dst[(int64_t)q * a.P + n] = (C)g_in[(n - n0) * Q7 + q];
}

// mass, min, max, first and second moments, non-finite count (FP64, fixed reduction order)
constexpr int ST_NDIAG = 10;   // mass -cmin cmax m1[3] m2[3] nan
constexpr uint32_t ST_MAX_MASK = (1u << 1) | (1u << 2);

template <typename C, int MODE>
__global__ void __launch_bounds__(BLOCK) st_diag_kernel(StArgs<C> a, const int32_t *__restrict__ supinv, int32_t ny,
                                                        int32_t nz, double *__restrict__ partial)
{
    const int64_t n = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    double v[ST_NDIAG];
#pragma unroll
    for (int k = 0; k < ST_NDIAG; ++k) v[k] = 0.0;
    v[1] = v[2] = -INFINITY;
    if (n < a.N) {
        double c;
        if (MODE == K_INIT) {
            c = (double)a.cinit[n];
        } else {
            C g[Q7];
            st_gather<C, MODE>(a, n, g);
            c = 0.0;
#pragma unroll
            for (int q = 0; q < Q7; ++q) c += (double)g[q];
        }
        const int32_t s = supinv[n];
        const int32_t nyz = ny * nz;
        const double x[3] = {(double)(s / nyz), (double)((s % nyz) / nz), (double)(s % nz)};
        if (isfinite(c)) {
            v[0] = c;
            v[1] = -c;
            v[2] = c;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                v[3 + d] = c * x[d];
                v[6 + d] = c * x[d] * x[d];
            }
        } else {
            v[9] = 1.0;
        }
    }
    const bool is_max[ST_NDIAG] = {false, true, true, false, false, false, false, false, false, false};
    block_reduce_store<ST_NDIAG>(v, is_max, partial + (int64_t)blockIdx.x * ST_NDIAG);
}

// ---------------------------------------------------------------- host side
namespace {

template <typename C>
StArgs<C> make_st_args(const void *src, void *dst, NbrDev nbr, const void *u, const void *cinit, int64_t N, int64_t P,
                       const lbm_st_params &p)
{
    StArgs<C> a;
    a.src = (const C *)src;
    a.dst = (C *)dst;
    a.nbr = nbr;
    a.u = (const C *)u;
    a.cinit = (const C *)cinit;
    a.N = N;
    a.P = P;
    const double cs2 = (1.0 - p.w0) / 3.0;
    a.omega = (C)(1.0 / (p.D / cs2 + 0.5));
    a.w0 = (C)p.w0;
    a.wq = (C)((1.0 - p.w0) / 6.0);
    a.inv_cs2 = (C)(1.0 / cs2);
    return a;
}

template <typename C>
hipError_t launch_st_t(const StArgs<C> &a, int mode, int64_t grid, const lbm_launch &ls, hipStream_t s)
{
    dim3 g((unsigned)grid), b(BLOCK);
    if (mode == K_INIT)
        st_step_kernel<C, K_INIT><<<g, b, 0, s>>>(a, ls.order);
    else if (mode == K_DIRECT)
        st_step_kernel<C, K_DIRECT><<<g, b, 0, s>>>(a, ls.order);
    else if (ls.block == 128)
        st_step_kernel<C, K_PULL, 128><<<dim3((unsigned)((a.N + 127) / 128)), dim3(128), 0, s>>>(a, ls.order);
    else if (ls.block == 512)
        st_step_kernel<C, K_PULL, 512><<<dim3((unsigned)((a.N + 511) / 512)), dim3(512), 0, s>>>(a, ls.order);
    else
        st_step_kernel<C, K_PULL><<<g, b, 0, s>>>(a, ls.order);
    return hipGetLastError();
}

int st_kmode(StateMode m) { return m == STATE_INIT ? K_INIT : (m == STATE_DIRECT ? K_DIRECT : K_PULL); }

}  // namespace

// Standalone: time the solute step's block sizes and tile orders (on a zero field, in the spare
// buffer: init or load_state sets the state).  Coupled: time the fused flow + solute launch on
// the flow's state, writing only both handles' spare buffers, so neither state changes.
int StSolver::autotune()
{
    if (N_ < TUNE_MIN_NODES) return LBM_OK;
    const int kind = flow_ ? -LBM_KERNEL_ST_STEP : LBM_KERNEL_ST_STEP;   // separate cache entries
    lbm_launch c;
    if (tune_cache_get(kind, prm_.precision, N_, c)) {
        (flow_ ? launch_ : launch_own_) = c;
        return LBM_OK;
    }
    for (void *b : pops_) LBM_HIP_TRY(hipMemsetAsync(b, 0, (size_t)Q7 * P_ * st_bytes_, stream_));
    lbm_launch best;
    if (flow_) {
        const double cs2 = (1.0 - prm_.w0) / 3.0;
        FusedSolute fs{pops_[0], pops_[1], 1.0 / (prm_.D / cs2 + 0.5), prm_.w0, (1.0 - prm_.w0) / 6.0, 1.0 / cs2, launch_};
        const void *state = flow_->pops(flow_->current());
        void *spare = flow_->pops(flow_->current() ^ 1);
        const auto cands = tune_candidates(launch_, {128, 256, 512}, {ORDER_HW, ORDER_RUNS8, ORDER_SLAB}, {0});
        LBM_HIP_TRY(tune_launch(stream_, cands,
                                [&](const lbm_launch &ls) {
                                    return flow_->launch_kernel(K_PULL, false, &fs, ls, state, spare, true);
                                },
                                best));
        launch_ = best;
    } else {
        const void *u = d_u_own_;
        const auto cands = tune_candidates(launch_own_, {128, 256, 512}, {ORDER_HW, ORDER_RUNS8, ORDER_SLAB}, {0});
        auto run = [&](const lbm_launch &ls) {
            return st_bytes_ == 8
                       ? launch_st_t(make_st_args<double>(pops_[0], pops_[1], nbr_, u, d_cinit_, N_, P_, prm_), K_PULL,
                                     grid_, ls, stream_)
                       : launch_st_t(make_st_args<float>(pops_[0], pops_[1], nbr_, u, d_cinit_, N_, P_, prm_), K_PULL,
                                     grid_, ls, stream_);
        };
        LBM_HIP_TRY(tune_launch(stream_, cands, run, best));
        launch_own_ = best;
    }
    tune_cache_put(kind, prm_.precision, N_, best);
    return LBM_OK;
}

int StSolver::get_launch(int kernel, lbm_launch *out) const
{
    if (kernel != LBM_KERNEL_ST_STEP || !out) return LBM_ERR_ARG;
    *out = flow_ ? launch_ : launch_own_;
    return LBM_OK;
}

int StSolver::set_launch(int kernel, const lbm_launch *in)
{
    if (kernel != LBM_KERNEL_ST_STEP) return fail(LBM_ERR_ARG, "a solute handle launches LBM_KERNEL_ST_STEP only");
    if (!in || !launch_valid(*in, {128, 256, 512}, {0}))
        return fail(LBM_ERR_ARG, "launch settings: block 128, 256 or 512, order 0-2, variant 0");
    lbm_launch l = *in;
    l.tuned = 0;
    l.time_us = 0.0;
    (flow_ ? launch_ : launch_own_) = l;
    return LBM_OK;
}

StSolver::~StSolver() { release(); }

void StSolver::release()
{
    (void)hipSetDevice(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (void *&b : pops_)
        if (b) (void)hipFree(b), b = nullptr;
    own_nbr_.release();
    void **bufs[] = {(void **)&own_supinv_, &d_u_own_, &d_cinit_, (void **)&d_partial_, (void **)&d_result_};
    for (void **b : bufs)
        if (*b) (void)hipFree(*b), *b = nullptr;
    if (h_result_) (void)hipHostFree(h_result_), h_result_ = nullptr;
    stage_.release();
    if (ev0_) (void)hipEventDestroy(ev0_), ev0_ = nullptr;
    if (ev1_) (void)hipEventDestroy(ev1_), ev1_ = nullptr;
    if (own_stream_) (void)hipStreamDestroy(own_stream_), own_stream_ = nullptr;
    stream_ = nullptr;
}

int StSolver::setup(const lbm_st_params &p)
{
    if (p.precision != LBM_FP64 && p.precision != LBM_FP32) return fail(LBM_ERR_ARG, "solute supports LBM_FP64 and LBM_FP32");
    if (!(p.D > 0.0)) return fail(LBM_ERR_ARG, "diffusion coefficient must be positive");
    if (!(p.w0 >= 0.0 && p.w0 < 1.0)) return fail(LBM_ERR_ARG, "rest weight must be in [0, 1)");
    st_bytes_ = p.precision == LBM_FP64 ? 8 : 4;
    grid_ = (N_ + BLOCK - 1) / BLOCK;
    LBM_HIP_TRY(hipEventCreate(&ev0_));
    LBM_HIP_TRY(hipEventCreate(&ev1_));
    for (void *&b : pops_) LBM_HIP_TRY(hipMalloc(&b, (size_t)Q7 * P_ * st_bytes_));
    LBM_HIP_TRY(hipMalloc(&d_cinit_, (size_t)P_ * st_bytes_));
    LBM_HIP_TRY(hipMalloc(&d_partial_, (size_t)grid_ * ST_NDIAG * sizeof(double)));
    LBM_HIP_TRY(hipMalloc(&d_result_, 16 * sizeof(double)));
    LBM_HIP_TRY(hipHostMalloc((void **)&h_result_, 16 * sizeof(double), hipHostMallocDefault));
    return LBM_OK;
}

int StSolver::create(const lbm_st_params &p, const int32_t *solid)
{
    prm_ = p;
    std::string err;
    int rc = build_sparse_geometry(p.nx, p.ny, p.nz, solid, geo_, err);
    if (rc != LBM_OK) return fail(rc, err);
    g_ = &geo_;
    N_ = geo_.N;
    P_ = geo_.P;
    device_ = p.device;
    LBM_HIP_TRY(hipSetDevice(device_));
    LBM_HIP_TRY(hipStreamCreateWithFlags(&own_stream_, hipStreamNonBlocking));
    stream_ = own_stream_;
    rc = setup(p);
    if (rc != LBM_OK) return rc;
    PackedNbr packed;
    rc = pack_neighbours(geo_, packed, err);
    if (rc != LBM_OK) return fail(rc, err);
    LBM_HIP_TRY(own_nbr_.upload(packed, P_, stream_));
    nbr_ = own_nbr_.view();
    LBM_HIP_TRY(hipMalloc(&own_supinv_, (size_t)P_ * sizeof(int32_t)));
    LBM_HIP_TRY(hipMemcpyAsync(own_supinv_, geo_.supinv.data(), geo_.supinv.size() * sizeof(int32_t),
                               hipMemcpyHostToDevice, stream_));
    d_supinv_ = own_supinv_;
    LBM_HIP_TRY(hipMalloc(&d_u_own_, (size_t)3 * P_ * st_bytes_));
    LBM_HIP_TRY(hipMemsetAsync(d_u_own_, 0, (size_t)3 * P_ * st_bytes_, stream_));
    LBM_HIP_TRY(hipStreamSynchronize(stream_));
    std::vector<uint32_t>().swap(geo_.nbr);
    rc = autotune();
    if (rc != LBM_OK) return rc;
    LBM_HIP_TRY(hipStreamSynchronize(stream_));
    return LBM_OK;
}

int StSolver::create_coupled(const lbm_st_params &p, SpSolver *flow)
{
    prm_ = p;
    flow_ = flow;
    g_ = &flow->geometry();
    N_ = g_->N;
    P_ = g_->P;
    prm_.nx = g_->nx;
    prm_.ny = g_->ny;
    prm_.nz = g_->nz;
    device_ = flow->device();
    if (p.device != device_) return fail(LBM_ERR_ARG, "coupled solute must live on the flow handle's device");
    const size_t want = p.precision == LBM_FP64 ? 8 : 4;
    if (flow->compute_bytes() != want)
        return fail(LBM_ERR_ARG, "coupled solute precision must match the flow handle's compute type");
    LBM_HIP_TRY(hipSetDevice(device_));
    stream_ = flow->stream();
    const int rc = setup(p);
    if (rc != LBM_OK) return rc;
    nbr_ = flow->nbr_view();
    d_supinv_ = flow->device_supinv();
    LBM_HIP_TRY(hipMalloc(&d_u_own_, (size_t)3 * P_ * st_bytes_));
    LBM_HIP_TRY(hipMemsetAsync(d_u_own_, 0, (size_t)3 * P_ * st_bytes_, stream_));
    const int rt = autotune();
    if (rt != LBM_OK) return rt;
    LBM_HIP_TRY(hipStreamSynchronize(stream_));
    return LBM_OK;
}

int StSolver::set_velocity(const double *u)
{
    if (flow_) return fail(LBM_ERR_STATE, "a coupled solute takes the flow handle's velocity");
    if (!u) return fail(LBM_ERR_ARG, "null velocity");
    return upload_velocity(u);
}

// node-major u[count][3] (double) to the 3 x P compute-type buffer d_u_own_
int StSolver::upload_velocity(const double *u)
{
    LBM_HIP_TRY(hipSetDevice(device_));
    const size_t cnt = (size_t)3 * P_;
    std::vector<unsigned char> t(cnt * st_bytes_, 0);
    for (int64_t n = 0; n < N_; ++n)
        for (int c = 0; c < 3; ++c) {
            const size_t i = (size_t)c * P_ + n;
            if (st_bytes_ == 8)
                ((double *)t.data())[i] = u[n * 3 + c];
            else
                ((float *)t.data())[i] = (float)u[n * 3 + c];
        }
    LBM_HIP_TRY(hipMemcpyAsync(d_u_own_, t.data(), t.size(), hipMemcpyHostToDevice, stream_));
    LBM_HIP_TRY(hipStreamSynchronize(stream_));
    return LBM_OK;
}

int StSolver::init(const double *conc_ijk)
{
    if (!conc_ijk) return fail(LBM_ERR_ARG, "null concentration field");
    LBM_HIP_TRY(hipSetDevice(device_));
    if (flow_) {
        // g = g_eq(C, u) with the velocity the flow's next step collides with; a copy of it
        // serves every read of the initial state (the flow's own buffer is written only by
        // its next step)
        if (!flow_->initialised()) return fail(LBM_ERR_STATE, "flow handle not initialised");
        std::vector<double> u((size_t)N_ * 3);
        const int rc = flow_->get_state(nullptr, u.data(), nullptr);
        if (rc != LBM_OK) return fail(rc, "flow velocity: " + flow_->error());
        const int rc2 = upload_velocity(u.data());
        if (rc2 != LBM_OK) return rc2;
    }
    if (st_bytes_ == 8) {
        std::vector<double> c((size_t)P_, 0.0);
        for (int64_t n = 0; n < N_; ++n) c[(size_t)n] = conc_ijk[g_->supinv[(size_t)n]];
        LBM_HIP_TRY(hipMemcpyAsync(d_cinit_, c.data(), c.size() * 8, hipMemcpyHostToDevice, stream_));
        LBM_HIP_TRY(hipStreamSynchronize(stream_));
    } else {
        std::vector<float> c((size_t)P_, 0.0f);
        for (int64_t n = 0; n < N_; ++n) c[(size_t)n] = (float)conc_ijk[g_->supinv[(size_t)n]];
        LBM_HIP_TRY(hipMemcpyAsync(d_cinit_, c.data(), c.size() * 4, hipMemcpyHostToDevice, stream_));
        LBM_HIP_TRY(hipStreamSynchronize(stream_));
    }
    mode_ = STATE_INIT;
    steps_ = 0;
    return LBM_OK;
}

int StSolver::load_state(const double *g)
{
    if (!g) return fail(LBM_ERR_ARG, "null state");
    LBM_HIP_TRY(hipSetDevice(device_));
    LBM_HIP_TRY(stage_.ensure());
    const int64_t chunk = stage_chunk(STAGE_BYTES, Q7);
    hipError_t e = hipSuccess;
    for (int64_t n0 = 0; n0 < N_ && e == hipSuccess; n0 += chunk) {
        const int64_t c = std::min(chunk, N_ - n0);
        double *dg = (double *)stage_.buf;
        e = hipMemcpyAsync(dg, g + n0 * Q7, (size_t)c * Q7 * sizeof(double), hipMemcpyHostToDevice, stream_);
        if (e != hipSuccess) break;
        dim3 gr((unsigned)((c + BLOCK - 1) / BLOCK)), b(BLOCK);
        if (st_bytes_ == 8)
            st_load_state_kernel<double><<<gr, b, 0, stream_>>>(
                make_st_args<double>(nullptr, pops_[cur_], nbr_, nullptr, nullptr, N_, P_, prm_), n0, n0 + c, dg);
        else
            st_load_state_kernel<float><<<gr, b, 0, stream_>>>(
                make_st_args<float>(nullptr, pops_[cur_], nbr_, nullptr, nullptr, N_, P_, prm_), n0, n0 + c, dg);
        e = hipGetLastError();
    }
    const hipError_t es = hipStreamSynchronize(stream_);
    LBM_HIP_TRY(e);
    LBM_HIP_TRY(es);
    mode_ = STATE_DIRECT;
    steps_ = 0;
    return LBM_OK;
}

int StSolver::launch_step(bool last)
{
    const int nxt = cur_ ^ 1;
    if (flow_ && mode_ == STATE_PULL && flow_->pull_state()) {
        // both states in pull form: one launch advances the flow and the solute; the flow keeps
        // its velocity for its Error diagnostic on the last step of a call, as its own step does
        const double cs2 = (1.0 - prm_.w0) / 3.0;
        FusedSolute fs{pops_[cur_], pops_[nxt], 1.0 / (prm_.D / cs2 + 0.5), prm_.w0, (1.0 - prm_.w0) / 6.0, 1.0 / cs2,
                       launch_};
        const int rc = flow_->step_fused(fs, last);
        if (rc != LBM_OK) return fail(rc, "coupled flow step failed: " + flow_->error());
        cur_ = nxt;
        ++steps_;
        return LBM_OK;
    }
    if (flow_) {
        const int rc = flow_->step_storing_velocity();
        if (rc != LBM_OK) return fail(rc, "coupled flow step failed: " + flow_->error());
    }
    const void *u = velocity();
    hipError_t e;
    if (st_bytes_ == 8)
        e = launch_st_t(make_st_args<double>(pops_[cur_], pops_[nxt], nbr_, u, d_cinit_, N_, P_, prm_), st_kmode(mode_),
                        grid_, launch_own_, stream_);
    else
        e = launch_st_t(make_st_args<float>(pops_[cur_], pops_[nxt], nbr_, u, d_cinit_, N_, P_, prm_), st_kmode(mode_),
                        grid_, launch_own_, stream_);
    LBM_HIP_TRY(e);
    cur_ = nxt;
    mode_ = STATE_PULL;
    ++steps_;
    return LBM_OK;
}

int StSolver::step(int nsteps, double *elapsed_ms)
{
    if (nsteps < 0) return fail(LBM_ERR_ARG, "negative step count");
    if (mode_ == STATE_EMPTY) return fail(LBM_ERR_STATE, "state not initialised");
    if (flow_ && !flow_->initialised()) return fail(LBM_ERR_STATE, "flow handle not initialised");
    LBM_HIP_TRY(hipSetDevice(device_));
    if (elapsed_ms) LBM_HIP_TRY(hipEventRecord(ev0_, stream_));
    for (int s = 0; s < nsteps; ++s) {
        const int rc = launch_step(s == nsteps - 1);
        if (rc != LBM_OK) return rc;
    }
    if (elapsed_ms) {
        LBM_HIP_TRY(hipEventRecord(ev1_, stream_));
        LBM_HIP_TRY(hipEventSynchronize(ev1_));
        float ms = 0.f;
        LBM_HIP_TRY(hipEventElapsedTime(&ms, ev0_, ev1_));
        *elapsed_ms = ms;
    }
    return LBM_OK;
}

int StSolver::get_state(double *g, double *conc)
{
    if (mode_ == STATE_EMPTY) return fail(LBM_ERR_STATE, "state not initialised");
    if (!g && !conc) return LBM_OK;
    LBM_HIP_TRY(hipSetDevice(device_));
    LBM_HIP_TRY(stage_.ensure());
    const void *u = velocity();
    const int mode = st_kmode(mode_);
    const int64_t chunk = stage_chunk(STAGE_BYTES, (g ? Q7 : 0) + (conc ? 1 : 0));
    hipError_t e = hipSuccess;
    for (int64_t n0 = 0; n0 < N_ && e == hipSuccess; n0 += chunk) {
        const int64_t c = std::min(chunk, N_ - n0);
        double *dg = g ? (double *)stage_.buf : nullptr;
        double *dc = conc ? (double *)stage_.buf + (g ? c * Q7 : 0) : nullptr;
        const int64_t n1 = n0 + c;
        dim3 gr((unsigned)((n1 - group_floor(n0) + BLOCK - 1) / BLOCK)), b(BLOCK);
        if (st_bytes_ == 8) {
            auto a = make_st_args<double>(pops_[cur_], nullptr, nbr_, u, d_cinit_, N_, P_, prm_);
            if (mode == K_INIT) st_get_state_kernel<double, K_INIT><<<gr, b, 0, stream_>>>(a, n0, n1, dg, dc);
            else if (mode == K_DIRECT) st_get_state_kernel<double, K_DIRECT><<<gr, b, 0, stream_>>>(a, n0, n1, dg, dc);
            else st_get_state_kernel<double, K_PULL><<<gr, b, 0, stream_>>>(a, n0, n1, dg, dc);
        } else {
            auto a = make_st_args<float>(pops_[cur_], nullptr, nbr_, u, d_cinit_, N_, P_, prm_);
            if (mode == K_INIT) st_get_state_kernel<float, K_INIT><<<gr, b, 0, stream_>>>(a, n0, n1, dg, dc);
            else if (mode == K_DIRECT) st_get_state_kernel<float, K_DIRECT><<<gr, b, 0, stream_>>>(a, n0, n1, dg, dc);
            else st_get_state_kernel<float, K_PULL><<<gr, b, 0, stream_>>>(a, n0, n1, dg, dc);
        }
        e = hipGetLastError();
        if (e == hipSuccess && g)
            e = hipMemcpyAsync(g + n0 * Q7, dg, (size_t)c * Q7 * sizeof(double), hipMemcpyDeviceToHost, stream_);
        if (e == hipSuccess && conc)
            e = hipMemcpyAsync(conc + n0, dc, (size_t)c * sizeof(double), hipMemcpyDeviceToHost, stream_);
    }
    const hipError_t es = hipStreamSynchronize(stream_);
    LBM_HIP_TRY(e);
    LBM_HIP_TRY(es);
    return LBM_OK;
}

int StSolver::diagnostics(lbm_st_diag *out)
{
    if (!out) return fail(LBM_ERR_ARG, "null output");
    if (mode_ == STATE_EMPTY) return fail(LBM_ERR_STATE, "state not initialised");
    LBM_HIP_TRY(hipSetDevice(device_));
    const void *u = velocity();
    dim3 gr((unsigned)grid_), b(BLOCK);
    const int mode = st_kmode(mode_);
    if (st_bytes_ == 8) {
        auto a = make_st_args<double>(pops_[cur_], nullptr, nbr_, u, d_cinit_, N_, P_, prm_);
        if (mode == K_INIT) st_diag_kernel<double, K_INIT><<<gr, b, 0, stream_>>>(a, d_supinv_, g_->ny, g_->nz, d_partial_);
        else if (mode == K_DIRECT) st_diag_kernel<double, K_DIRECT><<<gr, b, 0, stream_>>>(a, d_supinv_, g_->ny, g_->nz, d_partial_);
        else st_diag_kernel<double, K_PULL><<<gr, b, 0, stream_>>>(a, d_supinv_, g_->ny, g_->nz, d_partial_);
    } else {
        auto a = make_st_args<float>(pops_[cur_], nullptr, nbr_, u, d_cinit_, N_, P_, prm_);
        if (mode == K_INIT) st_diag_kernel<float, K_INIT><<<gr, b, 0, stream_>>>(a, d_supinv_, g_->ny, g_->nz, d_partial_);
        else if (mode == K_DIRECT) st_diag_kernel<float, K_DIRECT><<<gr, b, 0, stream_>>>(a, d_supinv_, g_->ny, g_->nz, d_partial_);
        else st_diag_kernel<float, K_PULL><<<gr, b, 0, stream_>>>(a, d_supinv_, g_->ny, g_->nz, d_partial_);
    }
    LBM_HIP_TRY(hipGetLastError());
    reduce_partials_kernel<ST_NDIAG><<<1, BLOCK, 0, stream_>>>(d_partial_, grid_, ST_MAX_MASK, d_result_);
    LBM_HIP_TRY(hipGetLastError());
    LBM_HIP_TRY(hipMemcpyAsync(h_result_, d_result_, ST_NDIAG * sizeof(double), hipMemcpyDeviceToHost, stream_));
    LBM_HIP_TRY(hipStreamSynchronize(stream_));
    const double *r = h_result_;
    out->mass = r[0];
    out->cmin = -r[1];
    out->cmax = r[2];
    for (int d = 0; d < 3; ++d) {
        out->moment1[d] = r[3 + d];
        out->moment2[d] = r[6 + d];
    }
    out->nan_count = r[9];
    return LBM_OK;
}

}  // namespace lbm
