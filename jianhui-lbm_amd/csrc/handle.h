// handle.h -- the opaque lbm_handle behind the C ABI, and the single-phase solver object.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "lbm_mi355x.h"

struct lbm_handle {
    enum Kind { SP = 1, CG = 2, SERVER = 3, ST = 4 };
    explicit lbm_handle(Kind k) : kind(k) {}
    virtual ~lbm_handle() = default;
    Kind kind;
    // the message of the last failed call; a server handle's calls may come from several threads,
    // so it is written and read under err_mu
    int fail(int code, const std::string &msg)
    {
        std::lock_guard<std::mutex> lk(err_mu);
        last_error = msg;
        return code;
    }
    std::string error() const
    {
        std::lock_guard<std::mutex> lk(err_mu);
        return last_error;
    }

private:
    mutable std::mutex err_mu;
    std::string last_error;
};

namespace lbm {

// message returned by lbm_last_error(NULL) after a failed create
void set_create_error(const std::string &msg);

// Sparse fluid-node structure on the host: the reference's numbering (SP:1172-1193) and, per
// node and moving direction q = 1..18, the upstream node x - e_q in pull form (BB_FLAG |
// self when x - e_q is solid).  Stored direction-major with row stride P.
struct SparseGeometry {
    int nx = 0, ny = 0, nz = 0;
    int64_t N = 0;   // fluid nodes
    int64_t P = 0;   // padded row stride (multiple of the block size)
    std::vector<int32_t> supinv;   // node -> i*ny*nz + j*nz + k   (size N)
    std::vector<uint32_t> nbr;     // 18 x P upstream indices
};

// Number fluid voxels and build the pull neighbour table.  Returns LBM_OK or LBM_ERR_ARG.
int build_sparse_geometry(int nx, int ny, int nz, const int32_t *solid_ijk, SparseGeometry &g,
                          std::string &err);

// The reference's x-slab split of the box over nranks MPI ranks (Parallelize_Geometry,
// SP:1085-1108): ends[r] is the exclusive end of rank r's range of node ids (the numbering is
// i-major, so each rank's slab is a range).
void reference_rank_ends(const SparseGeometry &g, int nranks, std::vector<int64_t> &ends);

// Whether some cluster of pore nodes joined by the 18 lattice links wraps round the periodic box
// along axis (0 = x, 1 = y, 2 = z).  Without such a cluster there is no net throughflow along
// the axis at steady state: the permeability is zero up to the lattice's spurious currents.
bool pores_wrap_along(const SparseGeometry &g, int axis);

// Round up to a multiple of the block size so every direction row starts 256-byte aligned.
inline int64_t padded_stride(int64_t n) { return ((n + 255) / 256) * 256; }

// Device form of the neighbour lists (decoded by load_nbr in nbrpack.h): per wavefront of 64
// node ids and direction, a 32-bit base and 16-bit residuals; waves whose spread does not fit
// keep their full words in a side table.
constexpr int HDR_WORDS = 20;                  // per wave: 18 bases, wide slot, pad
constexpr int HDR_WIDE = 18;
constexpr uint16_t RES_BB = 0xFFFF;            // residual of a bounce-back link
constexpr uint32_t HDR_COMPACT = 0xFFFFFFFFu;  // wide slot of a compact wave

struct PackedNbr {
    std::vector<uint16_t> res;    // 18 x P
    std::vector<uint32_t> hdr;    // (P / 64) x HDR_WORDS
    std::vector<uint32_t> wide;   // n_wide x 18 x 64
    int64_t n_wide = 0;
};

// Pack g.nbr, then decode every entry and compare with the full table (LBM_ERR_STATE if any
// differs, which would be a packing bug).
int pack_neighbours(const SparseGeometry &g, PackedNbr &out, std::string &err);

struct NbrDev {
    const uint16_t *__restrict__ res;   // 18 x P
    const uint32_t *__restrict__ hdr;   // (P / 64) x HDR_WORDS
    const uint32_t *__restrict__ wide;  // n_wide x 18 x 64
};

// Owner of the packed table in device memory.
struct DeviceNbr {
    uint16_t *res = nullptr;
    uint32_t *hdr = nullptr, *wide = nullptr;
    int64_t n_wide = 0;
    size_t bytes = 0;
    // allocate and queue the copies on s; p must stay alive until s has been synchronised
    hipError_t upload(const PackedNbr &p, int64_t P, hipStream_t s);
    void release();
    NbrDev view() const { return NbrDev{res, hdr, wide}; }
};

// Face boundaries (lbm_sp_set_boundaries).  Every fluid node on an active face gets one
// entry with the rule that the reference's face loops leave on it; a rule that reads the
// inward neighbour's pre-collision state names a slot of the saved-state table, which the
// step kernel fills for the neighbour nodes (19 deviations, rho - rho0, u) and the next step
// reads.  Per wave of 64 node ids, two bit masks (boundary lanes, saved lanes) and the index
// of the wave's first entry and first slot.
enum BcRule : int32_t { BC_EQ = 0, BC_EQU = 1, BC_MOM = 2, BC_NEQ = 3 };
struct BcEntry {
    int32_t rule;
    int32_t slot;      // saved-state slot of the inward neighbour (BC_EQU / BC_MOM / BC_NEQ)
    int32_t quirk;     // x- face of Sub_BC 1 / 3: u_z taken from the neighbour's u_y
    int32_t pad;
    double val;        // boundary density
    double uf[3];      // BC_EQ: the equilibrium velocity
};
struct BcWave {
    uint64_t b, s;          // lanes that are boundary nodes / whose state is saved
    uint32_t bbase, sbase;  // first entry / slot of the wave
};
constexpr int BC_ROWS = 23;   // saved state per slot: h[19], rho - rho0, u[3]

struct BoundaryTables {
    std::vector<BcEntry> entry;    // sorted by node id
    std::vector<BcWave> wave;      // P / 64
    std::vector<int32_t> saved;    // node id of each slot, sorted
};

// Resolve the faces of bc on the geometry (LBM_ERR_ARG for unsupported or malformed input).
int build_boundaries(const SparseGeometry &g, const lbm_sp_bc &bc, BoundaryTables &out, std::string &err);

// device view; S holds BC_ROWS x sp values of the compute type, one buffer read (the
// previous step's) and one written per launch
struct BcDev {
    const BcWave *__restrict__ wave;   // nullptr: no faces
    const BcEntry *__restrict__ entry;
    const void *s_prev;
    void *s_cur;
    int64_t sp;
};

enum StateMode { STATE_EMPTY = 0, STATE_INIT = 1, STATE_DIRECT = 2, STATE_PULL = 3 };

constexpr int BLOCK_DEFAULT = 256;   // workgroup size of every launch the autotuner has not changed

// FP64 state crosses between host and device in chunks through one bounded device buffer,
// allocated on a handle's first state transfer and kept until the handle is destroyed, so a
// checkpoint of a large state (f alone is 0.9 GB in FP64 for 5.9 M nodes) needs no full-size
// scratch.
constexpr size_t STAGE_BYTES = size_t(16) << 20;
struct Staging {
    void *buf = nullptr;
    hipError_t ensure()
    {
        return buf ? hipSuccess : hipMalloc(&buf, STAGE_BYTES);
    }
    void release()
    {
        if (buf) (void)hipFree(buf), buf = nullptr;
    }
};
// nodes per chunk when each node carries per_node doubles through a buffer of the given size
// (a multiple of the block size, at least one block)
inline int64_t stage_chunk(size_t bytes, int per_node)
{
    const int64_t c = (int64_t)(bytes / (sizeof(double) * (size_t)per_node)) / 256 * 256;
    return c > 256 ? c : 256;
}

// a coupled solute's buffers and rates for the fused flow + solute launch (compute type of the
// flow handle)
struct FusedSolute {
    const void *src;
    void *dst;
    double omega, w0, wq, inv_cs2;
    lbm_launch ls;   // launch settings of the fused step (the solute handle's)
};

struct DiagBox;

class SpSolver : public lbm_handle {
public:
    SpSolver() : lbm_handle(SP) {}
    ~SpSolver() override;

    int create(const lbm_sp_params &p, const int32_t *solid);
    int init(double rho, const double u0[3]);
    int load_state(const double *f, const double *u, const double *rho);
    int step(int nsteps, double *elapsed_ms);
    int get_state(double *f, double *u, double *rho);
    // get_state through a caller-chosen device buffer of the given size (no allocation)
    int get_state_via(double *f, double *u, double *rho, void *buf, size_t bytes);
    // get_state through the library's fixed device array (lbm_sp_get_state_noalloc)
    int get_state_noalloc(double *f, double *u, double *rho);
    int diagnostics(const int32_t *box, lbm_sp_diag *out);
    // Error over the reference's nranks x-slab ranks: err, u_max, sum of |u|
    int error_ranks(int nranks, double out[3]);
    // per cube of a cube grid: sum of u and node count (lbm_sp_cube_sums)
    int cube_sums(const int32_t origin[3], int32_t size, int32_t m, double *out);
    int set_boundaries(const lbm_sp_bc *bc);
    int set_power_law(double m, double n);
    int get_viscosity(double *nu);
    int get_dense(double *u_ijk, double *rho_ijk);
    int get_launch(int kernel, lbm_launch *out) const;
    int set_launch(int kernel, const lbm_launch *in);

    // one step launch from buffer src to dst (bare: without the handle's faces and power law,
    // and without storing u -- the timing launches of the autotuners, which leave the state alone)
    hipError_t launch_kernel(int mode, bool store_u, const FusedSolute *fs, const lbm_launch &ls, const void *src,
                             void *dst, bool bare);
    void *pops(int k) const { return pops_[k]; }
    int current() const { return cur_; }

    int64_t count() const { return geo_.N; }
    int64_t steps_done() const { return steps_; }

    // for a coupled passive solute: one step that also stores its velocity, and the shared
    // device geometry
    int step_storing_velocity();
    // one flow step that also advances a coupled solute in the same launch (pull state only)
    int step_fused(const FusedSolute &fs, bool store_u);
    bool pull_state() const { return mode_ == STATE_PULL; }
    const void *device_velocity() const { return d_u_; }
    NbrDev nbr_view() const { return nbr_.view(); }
    const SparseGeometry &geometry() const { return geo_; }
    const int32_t *device_supinv() const { return d_supinv_; }
    hipStream_t stream() const { return stream_; }
    int device() const { return device_; }
    size_t compute_bytes() const { return ct_bytes_; }
    bool initialised() const { return mode_ != STATE_EMPTY; }

private:
    int launch_step(bool store_u, const FusedSolute *fs = nullptr);
    int autotune();
    int release();
    void release_boundaries();
    BcDev bc_view(bool for_step) const;
    int reset_viscosity();
    int reduce_range(const DiagBox &db, int64_t n_lo, int64_t n_hi);
    hipError_t state_to_device(int64_t n0, int64_t n1, const int32_t *supinv, int64_t v0, double *df, double *du,
                               double *dr);

    lbm_sp_params prm_{};
    SparseGeometry geo_;
    int device_ = 0;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    double sv_ = 1.0, sq_ = 1.0;
    float scale_ = 1.0f, inv_scale_ = 1.0f;
    size_t st_bytes_ = 8, ct_bytes_ = 8;
    void *pops_[2] = {nullptr, nullptr};
    int cur_ = 0;
    DeviceNbr nbr_;
    int32_t *d_supinv_ = nullptr;
    void *d_u_ = nullptr;     // 3 x P compute-type velocity of the previous step
    void *d_rho_ = nullptr;   // P compute-type density (loaded states only)
    double *d_partial_ = nullptr;
    double *d_result_ = nullptr;
    double *h_result_ = nullptr;   // pinned
    int64_t grid_ = 0;
    StateMode mode_ = STATE_EMPTY;
    double rho_i_ = 1.0, u_i_[3] = {0, 0, 0};
    int64_t steps_ = 0;
    // face boundaries (none unless set_boundaries was called)
    BcWave *d_bc_wave_ = nullptr;
    BcEntry *d_bc_entry_ = nullptr;
    void *d_bc_s_[2] = {nullptr, nullptr};
    int64_t bc_sp_ = 0;
    int bc_cur_ = 0;    // the saved-state buffer the next launch reads
    // power-law collision (none unless set_power_law was called)
    void *d_nu_vn_ = nullptr;   // P compute type
    double pl_m_ = 0.0, pl_n_ = 0.0;
    Staging stage_;   // host transfers of the FP64 state
    lbm_launch launch_{BLOCK_DEFAULT, 1, 0, 0, 0.0};   // the step's launch settings (runs of 8 tiles)
    // node lists of cube grids (cube_sums): per cube, entries (64-node group, lane mask) of its
    // nodes, cube c's entries at off[c] .. off[c+1]
    struct CubeGrid {
        int32_t key[5] = {0, 0, 0, 0, 0};   // origin, size, m
        int64_t ncubes = 0;
        uint32_t *off = nullptr, *grp = nullptr;
        uint64_t *mask = nullptr;
        double *sums = nullptr;             // 4 x ncubes
        uint64_t used = 0;                  // last use, for replacing the oldest
    };
    std::vector<CubeGrid> cube_grids_;
    uint64_t cube_clock_ = 0;
    int build_cube_grid(const int32_t origin[3], int32_t size, int32_t m, CubeGrid &g);
    static void free_cube_grid(CubeGrid &g);
};

class CgSolver : public lbm_handle {
public:
    CgSolver() : lbm_handle(CG) {}
    ~CgSolver() override;

    int create(const lbm_cg_params &p, const int32_t *solid);
    int init(const double *psi_ijk, const double u0[3]);
    int load_state(const double *f, const double *u, const double *rho, const double *psi, const double *rho_r,
                   const double *rho_b);
    int step(int nsteps, double *elapsed_ms);
    int set_body_force(const double g[3]);
    int get_state(double *f, double *u, double *rho, double *psi, double *rho_r, double *rho_b);
    int diagnostics(const int32_t *box, lbm_cg_diag *out);
    int get_launch(int kernel, lbm_launch *out) const;
    int set_launch(int kernel, const lbm_launch *in);

    int64_t count() const { return geo_.N; }
    int64_t steps_done() const { return steps_; }

private:
    int launch_step(bool store_u);
    hipError_t launch_collide(int mode, bool store_u, const lbm_launch &ls, const void *src, void *dst);
    hipError_t launch_colour(const lbm_launch &ls);
    int autotune();
    void release();

    lbm_cg_params prm_{};
    SparseGeometry geo_;
    int device_ = 0;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    double visc_[8] = {0};   // wl, wg, lg0, l1, l2, g1, g2, delta (CG:2368-2377)
    size_t ct_bytes_ = 8;
    void *pops_[2] = {nullptr, nullptr};
    int cur_ = 0;
    DeviceNbr nbr_;
    int32_t *d_supinv_ = nullptr;
    void *d_col_ = nullptr;   // 2 x P: rho_r, rho_b after colour streaming
    void *d_psi_ = nullptr;   // P: psi = (rho_r - rho_b)/(rho_r + rho_b)
    void *d_cs_ = nullptr;    // 8 x P (AoS per node): rho_r, rho_b, u, n of the last collision
    void *d_cwin_ = nullptr;  // FP32: per block, the runs of upstream colour states staged in LDS
    int cwin_slots_ = 0;      // LDS states per block of the staged colour kernel (0: gathers)
    int64_t cw_grid_ = 0;     // blocks of the staged kernels (512 nodes each)
    double cwin_fast_ = 0;    // fraction of the staged blocks that read without a search
    lbm_launch launch_collide_{BLOCK_DEFAULT, 1, 0, 0, 0.0};   // runs of 8 tiles
    lbm_launch launch_colour_{BLOCK_DEFAULT, 2, 0, 0, 0.0};    // gathers, one slab per XCD
    void *d_u_ = nullptr;     // 3 x P: u of the previous step
    void *d_rho_ = nullptr;   // P: rho of a loaded checkpoint
    double *d_partial_ = nullptr;
    double *d_result_ = nullptr;
    double *h_result_ = nullptr;
    int64_t grid_ = 0;
    StateMode mode_ = STATE_EMPTY;
    double u_i_[3] = {0, 0, 0};
    int64_t steps_ = 0;
    Staging stage_;
};

class StSolver : public lbm_handle {
public:
    StSolver() : lbm_handle(ST) {}
    ~StSolver() override;

    int create(const lbm_st_params &p, const int32_t *solid);
    int create_coupled(const lbm_st_params &p, SpSolver *flow);
    int set_velocity(const double *u);
    int init(const double *conc_ijk);
    int load_state(const double *g);
    int step(int nsteps, double *elapsed_ms);
    int get_state(double *g, double *conc);
    int diagnostics(lbm_st_diag *out);
    int get_launch(int kernel, lbm_launch *out) const;
    int set_launch(int kernel, const lbm_launch *in);
    int64_t count() const { return N_; }

private:
    int setup(const lbm_st_params &p);
    int autotune();
    int upload_velocity(const double *u);
    int launch_step(bool last);
    // the velocity a launch reads: the frozen (or, coupled, initial) field until the first
    // coupled step, then the one the flow step just stored
    const void *velocity() const { return (flow_ && mode_ != STATE_INIT) ? flow_->device_velocity() : d_u_own_; }
    void release();

    lbm_st_params prm_{};
    SpSolver *flow_ = nullptr;       // coupled: the flow handle (not owned)
    SparseGeometry geo_;             // standalone geometry (host); coupled: the flow's
    const SparseGeometry *g_ = nullptr;
    DeviceNbr own_nbr_;
    NbrDev nbr_{};
    int32_t *own_supinv_ = nullptr;
    const int32_t *d_supinv_ = nullptr;
    int64_t N_ = 0, P_ = 0, grid_ = 0;
    int device_ = 0;
    hipStream_t own_stream_ = nullptr, stream_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    size_t st_bytes_ = 8;
    void *pops_[2] = {nullptr, nullptr};
    int cur_ = 0;
    void *d_u_own_ = nullptr;        // frozen velocity (coupled: the flow's at init), 3 x P compute type
    void *d_cinit_ = nullptr;        // P compute type: concentration for the equilibrium start
    double *d_partial_ = nullptr, *d_result_ = nullptr, *h_result_ = nullptr;
    StateMode mode_ = STATE_EMPTY;
    int64_t steps_ = 0;
    Staging stage_;
    lbm_launch launch_{BLOCK_DEFAULT, 0, 0, 0, 0.0};       // coupled: the fused flow + solute step
    lbm_launch launch_own_{BLOCK_DEFAULT, 0, 0, 0, 0.0};   // standalone solute step
};

}  // namespace lbm
