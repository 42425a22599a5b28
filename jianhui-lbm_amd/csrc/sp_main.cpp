// sp_main.cpp -- lbm_sp: drop-in single-phase driver for SINGLE_PHASE_MPI_Software_Spars.
//
//   lbm_sp INPUT.dat [checkpoint_prefix] [--precision fp64|fp32|fp16] [--device N] [--ranks K]
//
// Reads the reference's INPUT.dat and geometry file, runs the time loop of SP:662-841 on one
// GPU through the C ABI, and writes the reference's result files under the output prefix:
// Results.txt, Permeability_error_local_Perm.txt (wr_per = 1), bodyforce.txt,
// Velocity_ave_max.txt, General_Results.txt, and the checkpoint files of Backup (SP:4507).
// Pressure and velocity faces (INPUT lines 7-12, Sub_BC) run on the device
// (lbm_sp_set_boundaries); the library refuses the velocity faces of Sub_BC 2 and 3.  The
// power-law collision (nnf_module = 1) runs on the device too (lbm_sp_set_power_law), and so do
// the field writers and the local and group permeabilities (loc_perm = 1, gperm > 0).  Lattice
// velocity scaling, outside the hot path, is refused, never silently approximated.
//
// --ranks K prints Error's velocity change and maximum speed (Results.txt, Velocity_ave_max.txt)
// as the reference computes them on K MPI ranks, each rank over its own x-slab without its last
// node (SP:3651-3685; lbm_sp_error_ranks), so that the files equal those of an `mpirun -np K` run
// of the reference.  The default, 1, is the single-domain value.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "host_io.h"
#include "lbm_mi355x.h"

using lbmhost::SpInput;

namespace {

int die(const std::string &msg)
{
    std::cerr << "lbm_sp: " << msg << std::endl;
    return 1;
}

int check(int rc, lbm_handle *h, const char *what)
{
    if (rc != LBM_OK) {
        const char *m = lbm_last_error(h);
        std::cerr << "lbm_sp: " << what << " failed (" << rc << "): " << (m ? m : "") << std::endl;
    }
    return rc;
}

void hms(double s, int &h, int &m, int &sec)
{
    h = (int)(s / 3600);
    m = (int)((s - h * 3600) / 60);
    sec = (int)(s - (h * 3600 + m * 60));
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::cerr << "usage: lbm_sp INPUT.dat [checkpoint_prefix] [--precision fp64|fp32|fp16] [--device N] [--ranks K]\n";
        return 2;
    }
    std::string input = argv[1], pfix = "./";
    int precision = LBM_FP64, device = 0, ranks = 1;
    for (int a = 2; a < argc; ++a) {
        if (!std::strcmp(argv[a], "--precision") && a + 1 < argc) {
            const std::string p = argv[++a];
            precision = p == "fp32" ? LBM_FP32 : (p == "fp16" ? LBM_FP16 : LBM_FP64);
        } else if (!std::strcmp(argv[a], "--device") && a + 1 < argc) {
            device = std::atoi(argv[++a]);
        } else if (!std::strcmp(argv[a], "--ranks") && a + 1 < argc) {
            ranks = std::atoi(argv[++a]);
            if (ranks < 1) return die("--ranks needs a positive rank count");
        } else {
            pfix = argv[a];
        }
    }
    SpInput in;
    std::string err;
    if (!lbmhost::read_sp_input(input, in, err)) return die(err);
    if (in.nnf_module != 0 && in.nnf_module != 1) return die("nnf_module must be 0 or 1");
    if ((int)in.lattice[0] == 1 && (in.lattice[1] != in.lattice[2]))
        return die("self-defined lattice velocity (lattice_v = 1) is not part of this build");
    if (in.loc_perm == 1 && in.nx < 2) return die("local permeability (loc_perm = 1) needs at least two x-planes");
    const int nx = in.nx, ny = in.ny, nz = in.nz;
    std::vector<int32_t> solid;
    if (!lbmhost::read_geometry(in.filename, nx, ny, nz, in.decbin, solid, err)) return die(err);
    int box[6];
    bool partial = false;
    lbmhost::perm_box(in.par_per, in.per_x, in.per_y, in.per_z, nx, ny, nz, box, partial);
    const double porosity = lbmhost::box_porosity(solid, nx, ny, nz, box);
    const double vbox = (double)(box[1] - box[0] + 1) * (box[3] - box[2] + 1) * (box[5] - box[4] + 1);
    const double vfull = (double)nx * ny * nz;

    lbm_sp_params p{};
    p.nx = nx;
    p.ny = ny;
    p.nz = nz;
    p.nu = in.in_vis;
    for (int a = 0; a < 3; ++a) p.g[a] = in.g[a];
    p.rho0 = 1.0;
    p.precision = precision;
    p.device = device;
    lbm_handle *h = nullptr;
    if (check(lbm_sp_create(&p, solid.data(), &h), nullptr, "lbm_sp_create")) return 1;
    // pressure / velocity faces (INPUT lines 7-12 and Sub_BC; applied whenever a face flag is on,
    // SP:676-682): pre[axis] = {on+, p+, on-, p-}
    lbm_sp_bc bc{};
    bc.sub_bc = in.Sub_BC;
    bool faces = false;
    for (int a = 0; a < 3; ++a) {
        bc.pre[2 * a] = (int32_t)in.pre[a][0];
        bc.p[2 * a] = in.pre[a][1];
        bc.pre[2 * a + 1] = (int32_t)in.pre[a][2];
        bc.p[2 * a + 1] = in.pre[a][3];
        bc.vel[2 * a] = (int32_t)in.vel[a][0];
        bc.v[2 * a] = in.vel[a][1];
        bc.vel[2 * a + 1] = (int32_t)in.vel[a][2];
        bc.v[2 * a + 1] = in.vel[a][3];
        faces |= bc.pre[2 * a] == 1 || bc.pre[2 * a + 1] == 1 || bc.vel[2 * a] == 1 || bc.vel[2 * a + 1] == 1;
    }
    if (faces && check(lbm_sp_set_boundaries(h, &bc), h, "lbm_sp_set_boundaries")) return 1;
    // power-law fluid: collision_nnf in place of collision (SP:665-668)
    if (in.nnf_module == 1 && check(lbm_sp_set_power_law(h, in.nnf_m, in.nnf_n), h, "lbm_sp_set_power_law")) return 1;
    // dp of Comput_Perm (SP:4600-4614): the pressure drop along PerDir when in_BC = 1
    double dp = 0.0;
    if (in.in_BC != 0) {
        const int ax = (in.PerDir >= 1 && in.PerDir <= 3) ? in.PerDir - 1 : 0;
        const int len[3] = {nx, ny, nz};
        dp = std::fabs(in.pre[ax][1] - in.pre[ax][3]) * (1.0 / 3.0) / len[ax] / 1.0;
    }
    int64_t count = 0;
    lbm_sp_count(h, &count);
    const auto t_start = std::chrono::steady_clock::now();

    if (in.mode_backup_ini == 0) {
        if (check(lbm_sp_init(h, 1.0, in.ini_v), h, "lbm_sp_init")) return 1;
    } else {   // Backup_init (SP:4715-4828): files of label mode_backup_ini under the prefix
        std::vector<double> f((size_t)count * 19), u((size_t)count * 3), rho((size_t)count);
        if (!lbmhost::read_checkpoint(lbmhost::checkpoint_name(pfix, "rho", in.mode_backup_ini), rho.data(), count, 1, err) ||
            !lbmhost::read_checkpoint(lbmhost::checkpoint_name(pfix, "velocity", in.mode_backup_ini), u.data(), count, 3, err) ||
            !lbmhost::read_checkpoint(lbmhost::checkpoint_name(pfix, "f", in.mode_backup_ini), f.data(), count, 19, err))
            return die(err);
        if (check(lbm_sp_load_state(h, f.data(), u.data(), rho.data()), h, "lbm_sp_load_state")) return 1;
    }
    std::cout << "\nINITIALIZATION COMPLETED  (" << count << " fluid nodes, porosity " << porosity << ")\n\n";
    if ((in.freVe >= 0 || in.freDe >= 0) &&
        !lbmhost::write_vtk_geometry(in.outputfile + "LBM_Geometry.vtk", nx, ny, nz, solid.data(), err))   // SP:538-543
        return die(err);
    // field writers: the device scatters u / rho onto the box (lbm_sp_get_dense), the host
    // formats the reference's text in parallel
    auto secs = [](std::chrono::steady_clock::time_point t0) {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    };
    auto write_velocity = [&](long label) -> bool {
        const auto t0 = std::chrono::steady_clock::now();
        std::vector<double> ud((size_t)nx * ny * nz * 3);
        if (check(lbm_sp_get_dense(h, ud.data(), nullptr), h, "lbm_sp_get_dense")) return false;
        const double tg = secs(t0);
        const auto t1 = std::chrono::steady_clock::now();
        const std::string path = in.outputfile + "LBM_velocity_Vector_" + std::to_string(label) + ".vtk";
        const bool ok = lbmhost::write_vtk_velocity(path, nx, ny, nz, ud.data(), err) &&
                        lbmhost::write_velocity_profile("LBM_velocity_" + std::to_string(label) + ".out", nx, ny, nz,
                                                        solid.data(), ud.data(), err);
        std::cout << "[lbm_sp] velocity VTK " << label << ": device gather " << tg << " s, text " << secs(t1) << " s\n";
        return ok;
    };
    auto write_density = [&](long label) -> bool {
        const auto t0 = std::chrono::steady_clock::now();
        std::vector<double> rd((size_t)nx * ny * nz);
        if (check(lbm_sp_get_dense(h, nullptr, rd.data()), h, "lbm_sp_get_dense")) return false;
        const double tg = secs(t0);
        const auto t1 = std::chrono::steady_clock::now();
        const bool ok = lbmhost::write_vtk_density(in.outputfile + "LBM_Density_" + std::to_string(label) + ".vtk", nx,
                                                   ny, nz, rd.data(), err);
        std::cout << "[lbm_sp] density VTK " << label << ": device gather " << tg << " s, text " << secs(t1) << " s\n";
        return ok;
    };
    auto write_solute_velocity = [&](long label) -> bool {   // output_velocity_for_solute (SP:5093-5118)
        std::vector<double> u((size_t)count * 3);
        if (check(lbm_sp_get_state(h, nullptr, u.data(), nullptr), h, "lbm_sp_get_state")) return false;
        const std::string path = in.outputfile + "Velocity_for_solute_" + std::to_string(label) + ".0.bin";
        std::ofstream o(path, std::ios::binary);
        o.write((const char *)u.data(), (std::streamsize)(u.size() * sizeof(double)));
        if (!o) err = "write failed: " + path;
        return (bool)o;
    };

    const std::string out = in.outputfile;
    const std::string fResults = out + "Results.txt", fPerm = out + "Permeability_error_local_Perm.txt",
                      fBody = out + "bodyforce.txt", fVel = out + "Velocity_ave_max.txt";
    std::ofstream(fResults, std::ios::trunc).close();
    if (in.wr_per == 1) std::ofstream(fPerm, std::ios::trunc).close();
    std::ofstream(fBody, std::ios::out).close();
    std::ofstream(fVel, std::ios::out).close();

    auto backup = [&](long label) -> bool {
        std::vector<double> f((size_t)count * 19), u((size_t)count * 3), rho((size_t)count);
        if (check(lbm_sp_get_state(h, f.data(), u.data(), rho.data()), h, "lbm_sp_get_state")) return false;
        return lbmhost::write_checkpoint(lbmhost::checkpoint_name(out, "velocity", label), u.data(), count, 3, err) &&
               lbmhost::write_checkpoint(lbmhost::checkpoint_name(out, "rho", label), rho.data(), count, 1, err) &&
               lbmhost::write_checkpoint(lbmhost::checkpoint_name(out, "f", label), f.data(), count, 19, err);
    };

    const double rr = in.reso * in.reso * 1000;   // printed permeability unit (SP:742)
    double Permia[3] = {0, 0, 0}, error_perm = 0, u_max = 0, u_ave = 0, u_ave2 = -10000000, error = 0;
    double Permia_LOCAL[3] = {0, 0, 0};   // stays 0 unless loc_perm = 1, as in the reference
    double Re = 0;
    int U_max_ref = 0;
    const int dir = (in.PerDir >= 1 && in.PerDir <= 3) ? in.PerDir - 1 : 0;
    const double s_v = 1.0 / (in.in_vis * 3.0 + 0.5);
    int n = 0;
    long steps_done = 0;
    int th = 0, tm = 0, ts = 0, the = 0, tme = 0, tse = 0;
    while (n <= in.n_max) {
        // advance to the next diagnostics step (n % freRe == 0) or the end, in one call
        int next = (in.freRe > 0) ? ((n % in.freRe == 0) ? n : (n / in.freRe + 1) * in.freRe) : in.n_max;
        if (next > in.n_max) next = in.n_max;
        const int todo = next - n + 1;
        if (check(lbm_sp_step(h, todo), h, "lbm_sp_step")) return 1;
        steps_done += todo;
        n = next;
        if (in.freRe <= 0 || n % in.freRe != 0) {
            ++n;
            continue;
        }
        {   // rewrite the previous block (SP:692-707)
            std::ofstream fin(fResults, std::ios::out);
            fin << "The" << n - in.freRe << "th computation result:" << std::endl;
            fin << "The permiability is: " << Permia[0] * rr << ", " << Permia[1] * rr << ", " << Permia[2] * rr << std::endl;
            fin << "The relative error of permiability computing is: " << error_perm << std::endl;
            fin << "The LOCAL permiability is: " << Permia_LOCAL[0] * rr << ", " << Permia_LOCAL[1] * rr << ", "
                << Permia_LOCAL[2] * rr << std::endl;
            fin << "The Maximum velocity is: " << std::setprecision(6) << u_max << "   Re=" << Re
                << "     Courant Number=" << u_max << std::endl;
            fin << "The max relative error of velocity is: " << std::setiosflags(std::ios::scientific) << error << std::endl;
            fin << "Elapsed time is " << the << "h" << tme << "m" << tse << "s" << std::endl;
            fin << "The expected completion time is " << th << "h" << tm << "m" << ts << "s" << std::endl << std::endl;
        }
        // Error (SP:3625-3704) and Comput_Perm (SP:4565-4712) from the device reductions
        u_ave2 = u_ave;
        lbm_sp_diag dfull{}, dbox{};
        if (check(lbm_sp_diagnostics(h, nullptr, &dfull), h, "lbm_sp_diagnostics")) return 1;
        if (partial) {
            if (check(lbm_sp_diagnostics(h, box, &dbox), h, "lbm_sp_diagnostics")) return 1;
        } else {
            dbox = dfull;
        }
        error = dfull.err;
        u_max = dfull.umax;
        if (ranks > 1) {
            double er[3];
            if (check(lbm_sp_error_ranks(h, ranks, er), h, "lbm_sp_error_ranks")) return 1;
            error = er[0];
            u_max = er[1];
        }
        u_ave = dfull.sum_speed / (vfull * porosity);
        if (u_max >= 10.0) U_max_ref += 1;
        double Perm[3];
        for (int a = 0; a < 3; ++a) Perm[a] = dbox.Q[a] / vbox * in.in_vis / (in.g[a] + dp);
        error_perm = (Perm[dir] - Permia[dir]) / Permia[dir];
        for (int a = 0; a < 3; ++a) Permia[a] = Perm[a];
        if (in.loc_perm == 1) {
            // Comput_Perm_LOCAL (SP:4831-4883): the flux through the x-plane the last rank holds
            // as Solid[nx_l - 2], i.e. x = NX - 1, the plane before the last, over the per_*
            // cross-section of each direction
            const int32_t plane[6] = {nx - 2, nx - 2, 0, ny - 1, 0, nz - 1};
            lbm_sp_diag dl{};
            if (check(lbm_sp_diagnostics(h, plane, &dl), h, "lbm_sp_diagnostics")) return 1;
            const double ex = box[1] - box[0] + 1, ey = box[3] - box[2] + 1, ez = box[5] - box[4] + 1;
            const double area[3] = {ey * ez, ex * ez, ex * ey};
            for (int a = 0; a < 3; ++a) Permia_LOCAL[a] = dl.Q[a] / area[a] * in.in_vis / (in.g[a] + dp);
        }
        if (in.gperm > 0 && n % in.gperm == 0) {
            // Comput_Grop_Perm (SP:4886-5090): for each of the four cube grids, the permeability
            // along PerDir and the porosity of every cube (origin c + i*size, edge size, x then y
            // then z), one line per cube in Perm_<size>.output, rewritten at each call
            // (all cubes of a grid in one launch, lbm_sp_cube_sums)
            for (int gi = 0; gi < 4; ++gi) {
                const int L = in.size_gperm[gi], m = in.n_gperm[gi];
                std::ofstream fo(out + "Perm_" + std::to_string(L) + ".output", std::ios::out);
                if (m < 1 || L < 1) continue;
                const double vol = (double)L * L * L;
                std::vector<double> cs((size_t)4 * m * m * m);
                if (check(lbm_sp_cube_sums(h, in.c_gperm[gi], L, m, cs.data()), h, "lbm_sp_cube_sums")) return 1;
                for (size_t c = 0; c < (size_t)m * m * m; ++c) {
                    const double k = cs[4 * c + dir] / vol * in.in_vis / (in.g[dir] + dp);
                    fo << k * rr << " " << cs[4 * c + 3] / vol << std::endl;
                }
            }
        }
        {
            const double ax = dbox.Q[0] / (vbox * porosity), ay = dbox.Q[1] / (vbox * porosity), az = dbox.Q[2] / (vbox * porosity);
            u_ave = std::sqrt(ax * ax + ay * ay + az * az);
        }
        const double elaps = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
        const double remain = n > 0 ? (in.n_max - n) * (elaps / n) : 0.0;
        hms(remain, th, tm, ts);
        hms(elaps, the, tme, tse);
        Re = u_ave * ny / (1.0 / 3.0 * (1 / s_v - 0.5));
        {   // append this block (SP:725-757)
            std::ofstream fin(fResults, std::ios::app);
            fin << "The" << n << "th computation result:" << std::endl;
            fin << "The permiability is: " << Permia[0] * rr << ", " << Permia[1] * rr << ", " << Permia[2] * rr << std::endl;
            fin << "The LOCAL permiability is: " << Permia_LOCAL[0] * rr << ", " << Permia_LOCAL[1] * rr << ", "
                << Permia_LOCAL[2] * rr << std::endl;
            fin << "The relative error of permiability computing is: " << error_perm << std::endl;
            fin << "The Maximum velocity is: " << std::setprecision(6) << u_max << "   Re=" << Re
                << "     Courant Number=" << u_max << std::endl;
            fin << "The max relative error of velocity is: " << std::setiosflags(std::ios::scientific) << error << std::endl;
            fin << "Elapsed time is " << the << "h" << tme << "m" << tse << "s" << std::endl;
            fin << "The expected completion time is " << th << "h" << tm << "m" << ts << "s" << std::endl << std::endl;
        }
        if (in.wr_per == 1) {
            std::ofstream f(fPerm, std::ios::app);
            f << Permia[dir] * rr << " " << error_perm << " " << Permia_LOCAL[dir] * rr << std::endl;
        }
        std::ofstream(fBody, std::ios::app) << in.g[0] << std::endl;
        std::ofstream(fVel, std::ios::app) << u_ave << "  " << u_max << " " << error << std::endl;
        std::cout << "The" << n << "th computation result:\n"
                  << "The permiability is: " << Permia[0] * rr << ", " << Permia[1] * rr << ", " << Permia[2] * rr << "\n"
                  << "The relative error of permiability computing is: " << error_perm << "\n"
                  << "The Maximum velocity is: " << u_max << "   Re=" << Re << "\n"
                  << "The max relative error of uv is: " << error << "\n\n";
        bool stop = false;
        long nl = n;   // the label the outputs below see (n_max + 1 after the converged stop)
        if (in.Out_Mode == 1 && std::fabs((u_ave2 - u_ave) / u_ave2) < 1e-7) {   // SP:812-816
            if (!write_velocity(n)) return die(err);
            stop = true;
            nl = in.n_max + 1;
        }
        if (in.freDe > 0 && nl % in.freDe == 0 && !write_density(nl)) return die(err);   // SP:818-822
        if (in.freVe > 0 && nl % in.freVe == 0 && !write_velocity(nl)) return die(err);  // SP:824-828
        if (in.fre_backup > 0 && nl % in.fre_backup == 0 && nl > 0 && !backup(nl)) return die(err);
        if (in.vel_sol > 0 && nl % in.vel_sol == 0 && nl > 0 && !write_solute_velocity(nl)) return die(err);
        if (error != error) {
            std::cout << "PROGRAM STOP" << std::endl;
            stop = true;
        }
        if (U_max_ref >= 5) {
            std::cout << "PROGRAM STOP DUE TO HIGH VELOCITY" << std::endl;
            stop = true;
        }
        if (stop) break;
        ++n;
    }
    if (in.fre_backup >= 0 && !backup(in.n_max)) return die(err);
    const double elaps = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    hms(elaps, the, tme, tse);
    std::cout << "Elapsed time is " << the << "h" << tme << "m" << tse << "s" << std::endl;
    std::cout << "Fluid-node updates: " << (double)count * steps_done << "  (" << count * (double)steps_done / elaps / 1e6
              << " MLUPS including I/O)" << std::endl;
    {
        std::ofstream fin(fResults, std::ios::app);
        fin << "Elapsed time is " << the << "h" << tme << "m" << tse << "s" << std::endl;
    }
    {   // FOR SCRIPT RUN (SP:886-897)
        std::ofstream fin(out + "General_Results.txt", std::ios::app);
        fin << in.filename << "  " << Permia[0] * rr << " " << Permia[1] * rr << " " << Permia[2] * rr << std::endl << std::endl;
    }
    lbm_destroy(h);
    return 0;
}
