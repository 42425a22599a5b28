// cg_main.cpp -- lbm_cg: drop-in colour-gradient driver for Multi_Component_CG_MPI_SPARSE_Software.
//
//   lbm_cg INPUT_CG_SPARSE.dat [checkpoint_prefix] [--precision fp64|fp32] [--device N]
//
// Reads the reference's INPUT file, geometry and phase.dat (ini_Sat < 0), runs the time loop
// of CG:931-1182 on one GPU through the C ABI, and writes Results.txt,
// Relative_Permeability_Component{1,2}.txt, bodyforce.txt, Velocity_ave_max.txt and the
// checkpoints of Backup (CG:7020-7090).  The relative-permeability control modules (pressure
// or force schedules, capillary-pressure steps, least squares, imbibition/drainage control,
// input updates) are outside the hot path and are refused when switched on.
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

namespace {
int die(const std::string &msg)
{
    std::cerr << "lbm_cg: " << msg << std::endl;
    return 1;
}
int check(int rc, lbm_handle *h, const char *what)
{
    if (rc != LBM_OK) {
        const char *m = lbm_last_error(h);
        std::cerr << "lbm_cg: " << what << " failed (" << rc << "): " << (m ? m : "") << std::endl;
    }
    return rc;
}
}  // namespace

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::cerr << "usage: lbm_cg INPUT_CG_SPARSE.dat [checkpoint_prefix] [--precision fp64|fp32] [--device N]\n";
        return 2;
    }
    std::string input = argv[1], pfix = "./";
    int precision = LBM_FP64, device = 0;
    for (int a = 2; a < argc; ++a) {
        if (!std::strcmp(argv[a], "--precision") && a + 1 < argc) {
            const std::string p = argv[++a];
            precision = p == "fp32" ? LBM_FP32 : LBM_FP64;
        } else if (!std::strcmp(argv[a], "--device") && a + 1 < argc) {
            device = std::atoi(argv[++a]);
        } else {
            pfix = argv[a];
        }
    }
    lbmhost::CgInput in;
    std::string err;
    if (!lbmhost::read_cg_input(input, in, err)) return die(err);
    if (in.in_BC != 0 || in.in_psi_BC != 0) return die("pressure/velocity/psi boundaries are not part of this build");
    if (in.pressure_change > 0 || in.pressure_change2 > 0 || in.least_square > 0 || in.rel_perm_id_dir > 0 ||
        in.rel_perm_ift > 0 || in.input_dynamic > 0)
        return die("relative-permeability control modules are not part of this build");
    const int nx = in.nx, ny = in.ny, nz = in.nz;
    std::vector<int32_t> solid;
    if (!lbmhost::read_geometry(in.filename, nx, ny, nz, in.decbin, solid, err)) return die(err);
    int box[6];
    bool partial = false;
    lbmhost::perm_box(in.par_per, in.per_x, in.per_y, in.per_z, nx, ny, nz, box, partial);
    const double porosity = lbmhost::box_porosity(solid, nx, ny, nz, box);
    const double vbox = (double)(box[1] - box[0] + 1) * (box[3] - box[2] + 1) * (box[5] - box[4] + 1);
    const double vfull = (double)nx * ny * nz;

    lbm_cg_params p{};
    p.nx = nx;
    p.ny = ny;
    p.nz = nz;
    p.nu_l = in.niu_l;
    p.nu_g = in.niu_g;
    p.cos_theta = in.cos_theta;
    p.cap_a = in.CapA;
    // the stabiliser holds the force at zero until stab_time (CG:2536-2538, CG:934-935)
    double gs[3];
    for (int a = 0; a < 3; ++a) gs[a] = in.stab == 1 ? 0.0 : in.g[a];
    for (int a = 0; a < 3; ++a) p.g[a] = gs[a];
    p.bodyforce_apply = in.bodyforce_apply;
    p.rel_perm_psi = in.rel_perm_psi;
    p.rel_perm_psi2 = in.rel_perm_psi2;
    p.rho0 = 1.0;
    p.precision = precision;
    p.device = device;
    p.perm_dir = in.PerDir;
    lbm_handle *h = nullptr;
    if (check(lbm_cg_create(&p, solid.data(), &h), nullptr, "lbm_cg_create")) return 1;
    int64_t count = 0;
    lbm_cg_count(h, &count);
    const auto t_start = std::chrono::steady_clock::now();
    if (in.mode_backup_ini == 0) {
        if (in.ini_Sat >= 0) return die("random initial saturation (ini_Sat >= 0) is time-seeded in the reference; give a phase file");
        std::vector<double> psi;
        if (!lbmhost::read_phase(in.filenamepsi, nx, ny, nz, psi, err)) return die(err);
        if (check(lbm_cg_init(h, psi.data(), in.ini_v), h, "lbm_cg_init")) return 1;
    } else {
        const long L = in.mode_backup_ini;
        std::vector<double> f((size_t)count * 19), u((size_t)count * 3), rho((size_t)count), psi((size_t)count),
            rr((size_t)count), rb((size_t)count);
        if (!lbmhost::read_checkpoint(lbmhost::checkpoint_name(pfix, "rho", L), rho.data(), count, 1, err) ||
            !lbmhost::read_checkpoint(lbmhost::checkpoint_name(pfix, "velocity", L), u.data(), count, 3, err) ||
            !lbmhost::read_checkpoint(lbmhost::checkpoint_name(pfix, "f", L), f.data(), count, 19, err) ||
            !lbmhost::read_checkpoint(lbmhost::checkpoint_name(pfix, "psi", L), psi.data(), count, 1, err) ||
            !lbmhost::read_checkpoint(lbmhost::checkpoint_name(pfix, "rhor", L), rr.data(), count, 1, err) ||
            !lbmhost::read_checkpoint(lbmhost::checkpoint_name(pfix, "rhob", L), rb.data(), count, 1, err))
            return die(err);
        if (check(lbm_cg_load_state(h, f.data(), u.data(), rho.data(), psi.data(), rr.data(), rb.data()), h,
                  "lbm_cg_load_state"))
            return 1;
    }
    std::cout << "\nINITIALIZATION COMPLETED  (" << count << " fluid nodes, porosity " << porosity << ")\n\n";

    const std::string out = in.outputfile;
    const std::string fResults = out + "Results.txt", fRel1 = out + "Relative_Permeability_Component1.txt",
                      fRel2 = out + "Relative_Permeability_Component2.txt", fBody = out + "bodyforce.txt",
                      fVel = out + "Velocity_ave_max.txt";
    std::ofstream(fResults, std::ios::trunc).close();
    if (in.wr_per == 1) std::ofstream(fRel1, std::ios::trunc).close();
    std::ofstream(fBody, std::ios::out).close();
    std::ofstream(fRel2, std::ios::out).close();
    std::ofstream(fVel, std::ios::out).close();

    auto backup = [&](long label) -> bool {
        std::vector<double> f((size_t)count * 19), u((size_t)count * 3), rho((size_t)count), psi((size_t)count),
            rr((size_t)count), rb((size_t)count);
        if (check(lbm_cg_get_state(h, f.data(), u.data(), rho.data(), psi.data(), rr.data(), rb.data()), h, "get_state"))
            return false;
        return lbmhost::write_checkpoint(lbmhost::checkpoint_name(out, "velocity", label), u.data(), count, 3, err) &&
               lbmhost::write_checkpoint(lbmhost::checkpoint_name(out, "rho", label), rho.data(), count, 1, err) &&
               lbmhost::write_checkpoint(lbmhost::checkpoint_name(out, "psi", label), psi.data(), count, 1, err) &&
               lbmhost::write_checkpoint(lbmhost::checkpoint_name(out, "f", label), f.data(), count, 19, err) &&
               lbmhost::write_checkpoint(lbmhost::checkpoint_name(out, "rhor", label), rr.data(), count, 1, err) &&
               lbmhost::write_checkpoint(lbmhost::checkpoint_name(out, "rhob", label), rb.data(), count, 1, err);
    };

    const double rr_unit = in.reso * in.reso * 1000 / in.Permeability;   // relative permeability print (CG:1071)
    const int dir = (in.PerDir >= 1 && in.PerDir <= 3) ? in.PerDir - 1 : 0;
    double Per_l[3] = {0, 0, 0}, Per_g[3] = {0, 0, 0}, S_l = 0, u_max = 0, u_ave = 0, error = 0, error_Per = 0;
    int U_max_ref = 0;
    int n = 0;
    long steps_done = 0;
    while (n <= in.n_max) {
        int next = (in.freRe > 0) ? ((n % in.freRe == 0) ? n : (n / in.freRe + 1) * in.freRe) : in.n_max;
        if (next > in.n_max) next = in.n_max;
        // the stabiliser switches the force on at n == stab_time (CG:934-935)
        if (in.stab == 1 && in.stab_time >= n && in.stab_time <= next) {
            if (in.stab_time > n) {
                if (check(lbm_cg_step(h, in.stab_time - n), h, "lbm_cg_step")) return 1;
                steps_done += in.stab_time - n;
                n = in.stab_time;
            }
            for (int a = 0; a < 3; ++a) gs[a] = in.g[a];
            if (check(lbm_cg_set_body_force(h, gs), h, "lbm_cg_set_body_force")) return 1;
        }
        const int todo = next - n + 1;
        if (check(lbm_cg_step(h, todo), h, "lbm_cg_step")) return 1;
        steps_done += todo;
        n = next;
        if (in.freRe <= 0 || n % in.freRe != 0) {
            ++n;
            continue;
        }
        lbm_cg_diag dfull{}, dbox{};
        if (check(lbm_cg_diagnostics(h, nullptr, &dfull), h, "lbm_cg_diagnostics")) return 1;
        if (partial) {
            if (check(lbm_cg_diagnostics(h, box, &dbox), h, "lbm_cg_diagnostics")) return 1;
        } else {
            dbox = dfull;
        }
        // Error (CG:4973-5040)
        error = dfull.err;
        u_max = dfull.umax;
        u_ave = dfull.sum_speed / (vfull * porosity);
        if (u_max >= 10.0) U_max_ref += 1;
        // Comput_Perm (CG:6298-6480)
        double Pl[3], Pg[3];
        for (int a = 0; a < 3; ++a) {
            Pl[a] = dbox.Ql[a] / vbox * in.niu_l / (gs[a] + 0.0);
            Pg[a] = dbox.Qg[a] / vbox * in.niu_g / (gs[a] + 0.0);
        }
        const double vx = (dbox.Ql[0] + dbox.Qg[0]) / vbox, vy = (dbox.Ql[1] + dbox.Qg[1]) / vbox,
                     vz = (dbox.Ql[2] + dbox.Qg[2]) / vbox;
        const double Capillary = std::sqrt(vx * vx + vy * vy + vz * vz) / porosity * in.niu_l / in.CapA;
        u_ave = std::sqrt((vx / porosity) * (vx / porosity) + (vy / porosity) * (vy / porosity) + (vz / porosity) * (vz / porosity));
        error_Per = (Pl[dir] - Per_l[dir]) / Per_l[dir];
        for (int a = 0; a < 3; ++a) {
            Per_l[a] = Pl[a];
            Per_g[a] = Pg[a];
        }
        // Comput_Perm_LOCAL (CG:6214-6290, called at every check, CG:1009): the two phases'
        // fluxes through the x-plane the last rank holds as Solid[nx_l - 2] (x = NX - 1), over
        // the per_* cross-section of each direction
        double Pl_loc[3] = {0, 0, 0}, Pg_loc[3] = {0, 0, 0};
        if (nx >= 2) {
            const int32_t plane[6] = {nx - 2, nx - 2, 0, ny - 1, 0, nz - 1};
            lbm_cg_diag dl{};
            if (check(lbm_cg_diagnostics(h, plane, &dl), h, "lbm_cg_diagnostics")) return 1;
            const double ex = box[1] - box[0] + 1, ey = box[3] - box[2] + 1, ez = box[5] - box[4] + 1;
            const double area[3] = {ey * ez, ex * ez, ex * ey};
            for (int a = 0; a < 3; ++a) {
                Pl_loc[a] = dl.Ql[a] / area[a] * in.niu_l / (gs[a] + 0.0);
                Pg_loc[a] = dl.Qg[a] / area[a] * in.niu_g / (gs[a] + 0.0);
            }
        }
        // Comput_Saturation (CG:6604-6763)
        const double S_l_r = S_l;
        S_l = dbox.wet_count / (vbox * porosity);
        {
            std::ofstream fin(fResults, std::ios::app);
            fin << "The" << n << "th computation result:" << std::endl;
            fin << "The Maximum velocity is: " << std::setprecision(6) << u_max << "   Re_l=" << u_ave * ny / in.niu_l
                << "   Re_g=" << u_ave * ny / in.niu_g << std::endl;
            fin << "Courant Number=" << u_max << "\t Capillary Num=" << Capillary << std::endl;
            fin << "The max relative error of velocity is: " << std::setiosflags(std::ios::scientific) << error << std::endl;
            fin << "The relative permeability of component 1 is " << Per_l[0] * rr_unit << ", " << Per_l[1] * rr_unit << ", "
                << Per_l[2] * rr_unit << std::endl;
            fin << "The relative permeability of component 2 is " << Per_g[0] * rr_unit << ", " << Per_g[1] * rr_unit << ", "
                << Per_g[2] * rr_unit << std::endl;
            fin << "Satuation of Component 1: " << S_l << ", The satuation of Component 2: " << 1 - S_l << std::endl;
            fin << "The relative error of permiability computing is: " << error_Per << std::endl << std::endl;
        }
        if (in.wr_per == 1) {
            std::ofstream f1(fRel1, std::ios::app);
            // the pressures of the PerDir faces (CG:1092-1101); the last column and the
            // component-2 slope are the least-squares module's, 0 with it off
            const double pn = dir == 1 ? in.p_yn : (dir == 2 ? in.p_zn : in.p_xn);
            const double pp = dir == 1 ? in.p_yp : (dir == 2 ? in.p_zp : in.p_xp);
            f1 << Per_l[dir] * rr_unit << " " << S_l << " " << Pl_loc[dir] * rr_unit << " " << 1 - S_l << " " << pn << " "
               << pp << " " << gs[dir] << " " << 0 << std::endl;
            std::ofstream f2(fRel2, std::ios::app);
            f2 << Per_g[dir] * rr_unit << " " << 1 - S_l << " " << Pg_loc[dir] * rr_unit << " " << S_l << " " << 0 << " " << 0
               << " " << std::endl;
        }
        std::ofstream(fBody, std::ios::app) << S_l << " " << 1 - S_l << " " << std::fabs((S_l_r - S_l) / S_l) << " "
                                            << dbox.sd_l << " " << dbox.sd_g << std::endl;
        std::ofstream(fVel, std::ios::app) << u_ave << " " << u_max << "  " << error << std::endl;
        std::cout << "The" << n << "th computation result:  S_l = " << S_l << "  k_r1 = " << Per_l[dir] * rr_unit
                  << "  k_r2 = " << Per_g[dir] * rr_unit << "  u_max = " << u_max << "\n";
        bool stop = false;
        if (in.fre_backup > 0 && n % in.fre_backup == 0 && n > 0 && !backup(n)) return die(err);
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
    std::cout << "Elapsed time " << elaps << " s, " << count * (double)steps_done / elaps / 1e6 << " MLUPS including I/O\n";
    lbm_destroy(h);
    return 0;
}
