// host_io.h -- the reference's files on the host side: INPUT parameter files, geometry and
// phase files, checkpoints.  Used by the drop-in driver programs (lbm_sp, lbm_cg), which talk
// to the device only through the C ABI in lbm_mi355x.h.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace lbmhost {

// Single-phase INPUT.dat, the line map of SP:255-308 (grid sizes as counts)
struct SpInput {
    std::string filename;
    int nx = 0, ny = 0, nz = 0;
    int n_max = 0;
    double reso = 1.0;
    int in_BC = 0;
    double g[3] = {0, 0, 0};
    double pre[3][4] = {{0}};
    double vel[3][4] = {{0}};
    double in_vis = 0.1;
    double ini_v[3] = {0, 0, 0};
    int wr_per = 1, PerDir = 1, freRe = 100, Out_Mode = 0, freVe = -1, freDe = -1;
    double lattice[3] = {0, 1, 1};
    std::string outputfile = "./";
    int Sub_BC = 0, loc_perm = 0;
    int par_per[3] = {0, 0, 0};
    int per_x[2] = {0, 0}, per_y[2] = {0, 0}, per_z[2] = {0, 0};   // (p, n) as written
    int fre_backup = -1, mode_backup_ini = 0, vel_sol = -1;
    int gperm = 0;
    int n_gperm[4] = {1, 1, 1, 1}, size_gperm[4] = {8, 8, 8, 8};   // lines 38-39
    int c_gperm[4][3] = {{0}};                                      // lines 40-43: cube-grid origins
    int decbin = 0;
    int nnf_module = 0;
    double nnf_m = 0.0, nnf_n = 1.0;   // line 48, read when nnf_module = 1 (SP:307-308)
};

// Colour-gradient INPUT_CG_SPARSE.dat, the line map of CG:371-451
struct CgInput {
    std::string filename, filenamepsi;
    int nx = 0, ny = 0, nz = 0;
    int n_max = 0;
    double reso = 1.0;
    int in_BC = 0, in_psi_BC = 0;
    double g[3] = {0, 0, 0};
    double p_xp = 1.0, p_xn = 1.0;   // pressure-boundary values (lines 10-12), echoed in the rel-perm file
    double p_yp = 1.0, p_yn = 1.0, p_zp = 1.0, p_zn = 1.0;
    double niu_l = 0.1, niu_g = 0.1, cos_theta = 0, CapA = 0.01;
    double ini_v[3] = {0, 0, 0};
    double Permeability = 1.0;
    int wr_per = 1, PerDir = 1, freRe = 100, Out_Mode = 0, freVe = -1, freDe = -1, frePsi = -1;
    std::string outputfile = "./";
    int stab = 0, stab_time = 0;
    double ini_Sat = -1;
    int ini_buf = 0;
    double rel_perm_psi = 0.99, rel_perm_psi2 = 0.92;
    int par_per[3] = {0, 0, 0};
    int per_x[2] = {0, 0}, per_y[2] = {0, 0}, per_z[2] = {0, 0};
    int fre_backup = -1, mode_backup_ini = 0;
    int decbin = 0;
    int pressure_change = 0, bodyforce_apply = 0, pressure_change2 = 0, least_square = 0;
    int rel_perm_id_dir = 0, rel_perm_ift = 0, input_dynamic = 0;
};

bool read_sp_input(const std::string &path, SpInput &in, std::string &err);
bool read_cg_input(const std::string &path, CgInput &in, std::string &err);

// geometry: decbin = 1 -> raw int32 [i][j][k] (SP:1058-1065); 0 -> ASCII 0/1 tokens, x fastest,
// other tokens skipped (SP:1025-1036).  solid is returned [i][j][k], 0 = pore.
bool read_geometry(const std::string &path, int nx, int ny, int nz, int decbin, std::vector<int32_t> &solid,
                   std::string &err);
// phase field: ASCII floats, x fastest, read as float and clamped to [-1, 1] (CG:2114-2161)
bool read_phase(const std::string &path, int nx, int ny, int nz, std::vector<double> &psi, std::string &err);

// checkpoint files in the reference's layout: one record of padding, then count records
// (SP:4507-4562); name = prefix + "LBM_checkpoint_<what>_<label>.<rank>.bin_input"
std::string checkpoint_name(const std::string &prefix, const char *what, long label, int rank = 0);
bool write_checkpoint(const std::string &path, const double *data, int64_t count, int width, std::string &err);
bool read_checkpoint(const std::string &path, double *data, int64_t count, int width, std::string &err);

// The reference's legacy-VTK text files (Geometry_b SP:4168-4263, output_velocity_b SP:4265-4390,
// output_density_b SP:4405-4505): x fastest, values as its ostreams print them (six
// significant digits, "%g").  The text is formatted in parallel, one block of planes per
// thread, and written in order.  Dense arrays are [i][j][k] (u: [i][j][k][3]).
bool write_vtk_geometry(const std::string &path, int nx, int ny, int nz, const int32_t *solid_ijk, std::string &err);
bool write_vtk_velocity(const std::string &path, int nx, int ny, int nz, const double *u_ijk, std::string &err);
bool write_vtk_density(const std::string &path, int nx, int ny, int nz, const double *rho_ijk, std::string &err);
// LBM_velocity_<m>.out: u_x at (2, j, 1) for each j where (1, j, 1) is fluid, else 0
// (SP:4366-4375, solid (2, j, 1) reads the reference's padding record, written here as 0)
bool write_velocity_profile(const std::string &path, int nx, int ny, int nz, const int32_t *solid_ijk,
                            const double *u_ijk, std::string &err);

// porosity over the partial-permeability box (SP:1088) and that box as {xn,xp,yn,yp,zn,zp}
void perm_box(const int par_per[3], const int per_x[2], const int per_y[2], const int per_z[2], int nx, int ny, int nz,
              int box[6], bool &partial);
double box_porosity(const std::vector<int32_t> &solid, int nx, int ny, int nz, const int box[6]);

}  // namespace lbmhost
