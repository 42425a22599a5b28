// host_io.cpp -- reading and writing the reference's files (see host_io.h).
#include "host_io.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace lbmhost {
namespace {

struct Lines {
    std::vector<std::string> v;
    bool ok = false;
    explicit Lines(const std::string &path)
    {
        std::ifstream f(path);
        if (!f) return;
        std::string s;
        while (std::getline(f, s)) v.push_back(s);
        ok = true;
    }
    // k-th whitespace token (0-based) of 1-based line `line`
    bool tok(int line, int k, std::string &out) const
    {
        if (line < 1 || line > (int)v.size()) return false;
        std::istringstream is(v[line - 1]);
        std::string t;
        for (int i = 0; i <= k; ++i)
            if (!(is >> t)) return false;
        out = t;
        return true;
    }
};

// the reference reads with operator>>: an int slot takes the leading integer of the token
bool get_int(const Lines &L, int line, int k, int &v, std::string &err)
{
    std::string t;
    if (!L.tok(line, k, t)) {
        err = "line " + std::to_string(line) + ": missing value";
        return false;
    }
    v = (int)std::strtol(t.c_str(), nullptr, 10);
    return true;
}

bool get_dbl(const Lines &L, int line, int k, double &v, std::string &err)
{
    std::string t;
    if (!L.tok(line, k, t)) {
        err = "line " + std::to_string(line) + ": missing value";
        return false;
    }
    v = std::strtod(t.c_str(), nullptr);
    return true;
}

bool get_str(const Lines &L, int line, std::string &v, std::string &err)
{
    if (!L.tok(line, 0, v)) {
        err = "line " + std::to_string(line) + ": missing value";
        return false;
    }
    return true;
}

}  // namespace

bool read_sp_input(const std::string &path, SpInput &in, std::string &err)
{
    Lines L(path);
    if (!L.ok) {
        err = "cannot open " + path;
        return false;
    }
    bool ok = get_str(L, 2, in.filename, err) && get_int(L, 3, 0, in.nx, err) && get_int(L, 3, 1, in.ny, err) &&
              get_int(L, 3, 2, in.nz, err) && get_int(L, 4, 0, in.n_max, err) && get_dbl(L, 5, 0, in.reso, err) &&
              get_int(L, 6, 0, in.in_BC, err);
    for (int a = 0; ok && a < 3; ++a) ok = get_dbl(L, 7, a, in.g[a], err);
    for (int d = 0; ok && d < 3; ++d)
        for (int a = 0; ok && a < 4; ++a) ok = get_dbl(L, 8 + d, a, in.pre[d][a], err);
    for (int d = 0; ok && d < 3; ++d)
        for (int a = 0; ok && a < 4; ++a) ok = get_dbl(L, 11 + d, a, in.vel[d][a], err);
    ok = ok && get_dbl(L, 14, 0, in.in_vis, err);
    for (int a = 0; ok && a < 3; ++a) ok = get_dbl(L, 15, a, in.ini_v[a], err);
    ok = ok && get_int(L, 17, 0, in.wr_per, err) && get_int(L, 18, 0, in.PerDir, err) && get_int(L, 19, 0, in.freRe, err) &&
         get_int(L, 20, 0, in.Out_Mode, err) && get_int(L, 21, 0, in.freVe, err) && get_int(L, 22, 0, in.freDe, err);
    for (int a = 0; ok && a < 3; ++a) ok = get_dbl(L, 24, a, in.lattice[a], err);
    ok = ok && get_str(L, 25, in.outputfile, err) && get_int(L, 26, 0, in.Sub_BC, err) && get_int(L, 27, 0, in.loc_perm, err);
    for (int a = 0; ok && a < 3; ++a) ok = get_int(L, 28, a, in.par_per[a], err);
    ok = ok && get_int(L, 29, 0, in.per_x[0], err) && get_int(L, 29, 1, in.per_x[1], err) &&
         get_int(L, 30, 0, in.per_y[0], err) && get_int(L, 30, 1, in.per_y[1], err) &&
         get_int(L, 31, 0, in.per_z[0], err) && get_int(L, 31, 1, in.per_z[1], err) &&
         get_int(L, 33, 0, in.fre_backup, err) && get_int(L, 34, 0, in.mode_backup_ini, err) &&
         get_int(L, 35, 0, in.vel_sol, err) && get_int(L, 37, 0, in.gperm, err) && get_int(L, 45, 0, in.decbin, err) &&
         get_int(L, 47, 0, in.nnf_module, err);
    if (ok && in.nnf_module == 1) ok = get_dbl(L, 48, 0, in.nnf_m, err) && get_dbl(L, 48, 1, in.nnf_n, err);
    if (ok && in.gperm > 0) {   // group permeability grid (SP:298-303)
        for (int a = 0; ok && a < 4; ++a)
            ok = get_int(L, 38, a, in.n_gperm[a], err) && get_int(L, 39, a, in.size_gperm[a], err);
        for (int gi = 0; ok && gi < 4; ++gi)
            for (int a = 0; ok && a < 3; ++a) ok = get_int(L, 40 + gi, a, in.c_gperm[gi][a], err);
    }
    return ok;
}

bool read_cg_input(const std::string &path, CgInput &in, std::string &err)
{
    Lines L(path);
    if (!L.ok) {
        err = "cannot open " + path;
        return false;
    }
    bool ok = get_str(L, 2, in.filename, err) && get_str(L, 3, in.filenamepsi, err) && get_int(L, 4, 0, in.nx, err) &&
              get_int(L, 4, 1, in.ny, err) && get_int(L, 4, 2, in.nz, err) && get_int(L, 5, 0, in.n_max, err) &&
              get_dbl(L, 6, 0, in.reso, err) && get_int(L, 7, 0, in.in_BC, err) && get_int(L, 8, 0, in.in_psi_BC, err);
    for (int a = 0; ok && a < 3; ++a) ok = get_dbl(L, 9, a, in.g[a], err);
    ok = ok && get_dbl(L, 10, 1, in.p_xp, err) && get_dbl(L, 10, 3, in.p_xn, err) && get_dbl(L, 11, 1, in.p_yp, err) &&
         get_dbl(L, 11, 3, in.p_yn, err) && get_dbl(L, 12, 1, in.p_zp, err) && get_dbl(L, 12, 3, in.p_zn, err);
    ok = ok && get_dbl(L, 19, 0, in.niu_l, err) && get_dbl(L, 20, 0, in.niu_g, err) && get_dbl(L, 21, 0, in.cos_theta, err) &&
         get_dbl(L, 22, 0, in.CapA, err);
    for (int a = 0; ok && a < 3; ++a) ok = get_dbl(L, 23, a, in.ini_v[a], err);
    ok = ok && get_dbl(L, 24, 0, in.Permeability, err) && get_int(L, 26, 0, in.wr_per, err) &&
         get_int(L, 27, 0, in.PerDir, err) && get_int(L, 28, 0, in.freRe, err) && get_int(L, 29, 0, in.Out_Mode, err) &&
         get_int(L, 30, 0, in.freVe, err) && get_int(L, 31, 0, in.freDe, err) && get_int(L, 32, 0, in.frePsi, err) &&
         get_str(L, 35, in.outputfile, err) && get_int(L, 37, 0, in.stab, err) && get_int(L, 37, 1, in.stab_time, err) &&
         get_dbl(L, 38, 0, in.ini_Sat, err) && get_int(L, 39, 0, in.ini_buf, err) &&
         get_dbl(L, 40, 0, in.rel_perm_psi, err) && get_dbl(L, 40, 1, in.rel_perm_psi2, err);
    for (int a = 0; ok && a < 3; ++a) ok = get_int(L, 41, a, in.par_per[a], err);
    ok = ok && get_int(L, 42, 0, in.per_x[0], err) && get_int(L, 42, 1, in.per_x[1], err) &&
         get_int(L, 43, 0, in.per_y[0], err) && get_int(L, 43, 1, in.per_y[1], err) &&
         get_int(L, 44, 0, in.per_z[0], err) && get_int(L, 44, 1, in.per_z[1], err) &&
         get_int(L, 46, 0, in.fre_backup, err) && get_int(L, 47, 0, in.mode_backup_ini, err) &&
         get_int(L, 49, 0, in.decbin, err) && get_int(L, 51, 0, in.pressure_change, err) &&
         get_int(L, 57, 0, in.bodyforce_apply, err) && get_int(L, 59, 0, in.pressure_change2, err) &&
         get_int(L, 63, 0, in.least_square, err) && get_int(L, 65, 0, in.rel_perm_id_dir, err) &&
         get_int(L, 66, 0, in.rel_perm_ift, err) && get_int(L, 67, 0, in.input_dynamic, err);
    return ok;
}

bool read_geometry(const std::string &path, int nx, int ny, int nz, int decbin, std::vector<int32_t> &solid,
                   std::string &err)
{
    const size_t n = (size_t)nx * ny * nz;
    solid.assign(n, 0);
    if (decbin == 1) {
        std::ifstream f(path, std::ios::binary);
        if (!f) {
            err = "file open error on " + path;
            return false;
        }
        f.read(reinterpret_cast<char *>(solid.data()), (std::streamsize)(n * sizeof(int32_t)));
        if ((size_t)f.gcount() != n * sizeof(int32_t)) {
            err = path + ": shorter than nx*ny*nz int32 values";
            return false;
        }
        return true;
    }
    std::ifstream f(path);
    if (!f) {
        err = "the pore geometry file (" + path + ") does not exist";
        return false;
    }
    // tokens other than the numbers 0 and 1 are skipped, x runs fastest
    std::string t;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int v = -1;
                while (f >> t) {
                    char *end = nullptr;
                    const double d = std::strtod(t.c_str(), &end);
                    if (end != t.c_str() && *end == '\0' && (d == 0.0 || d == 1.0)) {
                        v = (int)d;
                        break;
                    }
                }
                if (v < 0) {
                    err = path + ": fewer than nx*ny*nz 0/1 values";
                    return false;
                }
                solid[((size_t)i * ny + j) * nz + k] = v;
            }
    return true;
}

bool read_phase(const std::string &path, int nx, int ny, int nz, std::vector<double> &psi, std::string &err)
{
    std::ifstream f(path);
    if (!f) {
        err = "the concentration file (" + path + ") does not exist";
        return false;
    }
    psi.assign((size_t)nx * ny * nz, 0.0);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                float v = 0.f;
                if (!(f >> v)) {
                    err = path + ": fewer than nx*ny*nz values";
                    return false;
                }
                v = v > 1.f ? 1.f : (v < -1.f ? -1.f : v);
                psi[((size_t)i * ny + j) * nz + k] = (double)v;
            }
    return true;
}

std::string checkpoint_name(const std::string &prefix, const char *what, long label, int rank)
{
    std::ostringstream s;
    s << prefix << "LBM_checkpoint_" << what << "_" << label << "." << rank << ".bin_input";
    return s.str();
}

bool write_checkpoint(const std::string &path, const double *data, int64_t count, int width, std::string &err)
{
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f) {
        err = "cannot write " + path;
        return false;
    }
    std::vector<double> pad((size_t)width, 0.0);   // the reference's unused record 0
    f.write(reinterpret_cast<const char *>(pad.data()), (std::streamsize)(width * sizeof(double)));
    f.write(reinterpret_cast<const char *>(data), (std::streamsize)(count * width * sizeof(double)));
    return (bool)f;
}

bool read_checkpoint(const std::string &path, double *data, int64_t count, int width, std::string &err)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = "file open error on " + path;
        return false;
    }
    f.seekg((std::streamoff)(width * sizeof(double)));
    f.read(reinterpret_cast<char *>(data), (std::streamsize)(count * width * sizeof(double)));
    if ((int64_t)f.gcount() != count * width * (int64_t)sizeof(double)) {
        err = path + ": shorter than expected";
        return false;
    }
    return true;
}

void perm_box(const int par_per[3], const int per_x[2], const int per_y[2], const int per_z[2], int nx, int ny, int nz,
              int box[6], bool &partial)
{
    // per_* are read as (p, n); a zero par_per flag means the whole axis (SP:994-999)
    box[0] = par_per[0] ? per_x[1] : 0;
    box[1] = par_per[0] ? per_x[0] : nx - 1;
    box[2] = par_per[1] ? per_y[1] : 0;
    box[3] = par_per[1] ? per_y[0] : ny - 1;
    box[4] = par_per[2] ? per_z[1] : 0;
    box[5] = par_per[2] ? per_z[0] : nz - 1;
    // the reference restricts the flux sum only when some par_per flag is exactly 1 (SP:4617)
    partial = (par_per[0] - 1) * (par_per[1] - 1) * (par_per[2] - 1) == 0;
}

double box_porosity(const std::vector<int32_t> &solid, int nx, int ny, int nz, const int box[6])
{
    int64_t pores = 0;
    for (int i = box[0]; i <= box[1]; ++i)
        for (int j = box[2]; j <= box[3]; ++j)
            for (int k = box[4]; k <= box[5]; ++k)
                if (solid[((size_t)i * ny + j) * nz + k] == 0) ++pores;
    const double vol = (double)(box[1] - box[0] + 1) * (box[3] - box[2] + 1) * (box[5] - box[4] + 1);
    (void)nx;
    return (double)pores / vol;
}

}  // namespace lbmhost

namespace lbmhost {

namespace {

const char *VTK_HEAD = "# vtk DataFile Version 2.0\n";

std::string vtk_header(const char *title, int nx, int ny, int nz, const char *data_line, bool lookup)
{
    std::ostringstream h;
    h << VTK_HEAD << title << "\nASCII\nDATASET STRUCTURED_POINTS\n";
    h << "DIMENSIONS         " << nx << "         " << ny << "         " << nz << "\n";
    h << "ORIGIN 0 0 0\nSPACING 1 1 1\n";
    h << "POINT_DATA     " << (long long)nx * ny * nz << "\n" << data_line << "\n";
    h << (lookup ? "LOOKUP_TABLE default\n" : "\n");
    return h.str();
}

// format the points of planes k in [k0, k1) (x fastest within a plane) with emit(i, j, k, buf)
template <typename Emit>
bool write_planes(const std::string &path, const std::string &head, int nx, int ny, int nz, Emit emit, std::string &err)
{
    std::FILE *fp = std::fopen(path.c_str(), "wb");
    if (!fp) {
        err = "cannot write " + path;
        return false;
    }
    std::fwrite(head.data(), 1, head.size(), fp);
    const int block = 8;   // planes per task
    std::vector<std::string> out((size_t)block);
    bool ok = true;
    for (int k0 = 0; k0 < nz && ok; k0 += block * 16) {
        const int k1 = std::min(nz, k0 + block * 16);
        const int ntask = (k1 - k0 + block - 1) / block;
        std::vector<std::string> parts((size_t)ntask);
#pragma omp parallel for schedule(dynamic, 1)
        for (int t = 0; t < ntask; ++t) {
            std::string &b = parts[(size_t)t];
            b.reserve((size_t)block * nx * ny * 16);
            char tmp[128];
            for (int k = k0 + t * block; k < std::min(k1, k0 + (t + 1) * block); ++k)
                for (int j = 0; j < ny; ++j)
                    for (int i = 0; i < nx; ++i) {
                        const int len = emit(i, j, k, tmp);
                        b.append(tmp, (size_t)len);
                    }
        }
        for (const auto &b : parts) ok = ok && std::fwrite(b.data(), 1, b.size(), fp) == b.size();
    }
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok) err = "write failed: " + path;
    return ok;
}

}  // namespace

bool write_vtk_geometry(const std::string &path, int nx, int ny, int nz, const int32_t *solid, std::string &err)
{
    const std::string head =
        vtk_header("J.Yang Lattice Boltzmann Simulation 3D Single Phase-Solid-Geometry", nx, ny, nz,
                   "SCALARS sample_scalars float", true);
    return write_planes(path, head, nx, ny, nz, [&](int i, int j, int k, char *buf) {
        const int64_t v = ((int64_t)i * ny + j) * nz + k;
        return std::snprintf(buf, 128, "%d\n", solid[v] != 0 ? 1 : 0);   // 1 = solid (SP:4189-4192)
    }, err);
}

bool write_vtk_velocity(const std::string &path, int nx, int ny, int nz, const double *u, std::string &err)
{
    const std::string head = vtk_header("J.Yang Lattice Boltzmann Simulation 3D Single Phase-Velocity", nx, ny, nz,
                                        "VECTORS sample_vectors double", false);
    return write_planes(path, head, nx, ny, nz, [&](int i, int j, int k, char *buf) {
        const double *p = u + (((int64_t)i * ny + j) * nz + k) * 3;
        return std::snprintf(buf, 128, "%g %g %g \n", p[0], p[1], p[2]);
    }, err);
}

bool write_vtk_density(const std::string &path, int nx, int ny, int nz, const double *rho, std::string &err)
{
    const std::string head =
        vtk_header("J.Yang Lattice Boltzmann Simulation 3D Single Phase-Solid-Density", nx, ny, nz,
                   "SCALARS sample_scalars float", true);
    return write_planes(path, head, nx, ny, nz, [&](int i, int j, int k, char *buf) {
        return std::snprintf(buf, 128, "%g\n", rho[((int64_t)i * ny + j) * nz + k]);
    }, err);
}

bool write_velocity_profile(const std::string &path, int nx, int ny, int nz, const int32_t *solid, const double *u,
                            std::string &err)
{
    std::FILE *fp = std::fopen(path.c_str(), "wb");
    if (!fp) {
        err = "cannot write " + path;
        return false;
    }
    for (int j = 0; j < ny; ++j) {
        double v = 0.0;
        if (nx > 2 && nz > 1 && solid[((int64_t)1 * ny + j) * nz + 1] == 0)
            v = u[(((int64_t)2 * ny + j) * nz + 1) * 3];
        std::fprintf(fp, "%g\n", v);
    }
    if (std::fclose(fp) != 0) {
        err = "write failed: " + path;
        return false;
    }
    return true;
}

}  // namespace lbmhost
