// momopt_gpu.cpp -- the input-deck driver of the Mn ("momopt") solver on one
// MI355X.  It keeps the reference driver's behaviour (src/main.cpp:45-251 with
// SOLVER momopt, Solver::initializeGrid main.cpp:257-372, MomOptSolver::init
// momopt_init.cpp:51-272, outputData momopt_output.cpp:21-77) on a single rank;
// the closure solves, flux and Heun update go through include/mnclosure.h.
//
// Usage: momopt_gpu [deck-path]      (default: ./input.deck, like the reference)
// Extra deck keys (optional): MOMOPT_GPU_PRECISION 0|1, MOMOPT_GPU_WARM_START 0|1|2,
// MOMOPT_GPU_LINE_SEARCH 0|1, MOMOPT_GPU_DEVICE n, and MOMOPT_TOL_GAMMA x: the
// supplementary stop test gamma <= x of optScaled (opt.cpp:288-293; the
// reference sets 1.1 at momopt_update.cpp:226 behind a comment, and its
// regression fixtures were written with it); absent or <= 0: off.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "input_deck.h"
#include "mnclosure.h"

namespace {

void fail(const char *msg, int line) {
    std::printf("%s:input error at line %d\n", msg, line);
    std::exit(1);
}

struct Grid {
    int nx, ny, NX, NY;
    double ax, bx, ay, by, dx, dy, init_x, init_y;
    std::vector<double> init, sigma_t, sigma_s;
};

// Solver::initializeGrid (main.cpp:257-372) for one rank.
void initialize_grid(Grid &g, int init_cond, double gaussian_sigma, double floor_v, double sigma) {
    g.NX = g.nx + 4;
    g.NY = g.ny + 4;
    const size_t nc = size_t(g.NX) * g.NY;
    g.init.assign(nc, 0.0);
    g.sigma_t.assign(nc, sigma);
    g.sigma_s.assign(nc, sigma);
    for (int i = 0; i < g.NX; ++i) {
        for (int j = 0; j < g.NY; ++j) {
            const double x = g.init_x + (i - 2) * g.dx + g.dx / 2.0;
            const double y = g.init_y + (j - 2) * g.dy + g.dy / 2.0;
            const size_t c = size_t(i) * g.NY + j;
            if (i < 2 || i > g.nx + 1 || j < 2 || j > g.ny + 1) {
                g.init[c] = floor_v;
                g.sigma_t[c] = g.sigma_s[c] = sigma;
            } else if (init_cond == 0 && gaussian_sigma == 0.0) {   // delta
                g.init[c] = floor_v;
                if (i == (g.nx + 2) / 2 && j == (g.ny + 2) / 2) g.init[c] = 1.0 / (g.dx * g.dy);
                g.sigma_t[c] = g.sigma_s[c] = sigma;
            } else if (init_cond == 0) {                              // Gaussian
                const double s2 = gaussian_sigma * gaussian_sigma;
                const double gf = 1.0 / (2.0 * M_PI * s2) * std::exp(-(x * x + y * y) / (2.0 * s2));
                g.init[c] = gf > floor_v ? gf : floor_v;
                g.sigma_t[c] = g.sigma_s[c] = sigma;
            } else if (init_cond == 1) {                              // lattice
                g.init[c] = floor_v;
                g.sigma_t[c] = g.sigma_s[c] = 1.0;
                static const double centres[11][2] = {{2, 2}, {2, 0}, {2, -2}, {1, 1}, {1, -1}, {0, -2},
                                                      {-1, 1}, {-1, -1}, {-2, 2}, {-2, 0}, {-2, -2}};
                for (const auto &cc : centres) {
                    if (std::fabs(x - cc[0]) < 0.5 && std::fabs(y - cc[1]) < 0.5) {
                        g.sigma_t[c] = 10.0;
                        g.sigma_s[c] = 0.0;
                    }
                }
            } else if (init_cond == 2) {                              // smooth periodic
                g.init[c] = 1 + std::sin(2 * M_PI * x) * std::cos(2 * M_PI * y);
                g.sigma_t[c] = g.sigma_s[c] = sigma;
            } else {
                g.init[c] = floor_v;
            }
        }
    }
}

void write_output(const Grid &g, mnc_ctx *ctx, int nm, double time, int opt_type) {
    char name[256];
    std::snprintf(name, sizeof name, "out_%.3f_%d.pn", time, 0);
    std::vector<double> mom(size_t(g.NX) * g.NY * nm);
    if (mnc_grid_get(ctx, mom.data(), nullptr) != MNC_OK) {
        std::printf("mnc_grid_get: %s\n", mnc_last_error(ctx));
        std::exit(1);
    }
    FILE *f = std::fopen(name, "wb");
    if (!f) { std::printf("cannot write %s\n", name); std::exit(1); }
    int64_t v = g.nx;
    std::fwrite(&v, sizeof v, 1, f);
    v = g.ny;
    std::fwrite(&v, sizeof v, 1, f);
    std::fwrite(&g.ax, sizeof(double), 1, f);
    std::fwrite(&g.ay, sizeof(double), 1, f);
    std::fwrite(&g.bx, sizeof(double), 1, f);
    std::fwrite(&g.by, sizeof(double), 1, f);
    v = nm;
    std::fwrite(&v, sizeof v, 1, f);
    for (int i = 2; i < 2 + g.nx; ++i)
        for (int j = 2; j < 2 + g.ny; ++j) std::fwrite(&mom[(size_t(i) * g.NY + j) * nm], sizeof(double), nm, f);
    std::fclose(f);
    std::snprintf(name, sizeof name, "out_%.3f_%d.opt", time, 0);
    mnc_stats st{};
    mnc_get_stats(ctx, &st);
    f = std::fopen(name, "w");
    if (!f) return;
    static const double reg[5] = {0.0, 1e-3, 1e-2, 1e-1, 5e-1};
    std::fprintf(f, "Max number of iterations: %d\n", st.max_iter);
    std::fprintf(f, "Mean number of iterations: %f\n",
                 st.num_dual_solves ? double(st.iter_sum) / double(st.num_dual_solves) : 0.0);
    std::fprintf(f, "Number of optimizations: %lld\n", st.num_dual_solves);
    std::fprintf(f, "Regularizations Histogram\n");
    for (int r = 0; r < 5; ++r) std::fprintf(f, "r = %.3e   %lld\n", reg[r], st.hist_reg[r]);
    std::fclose(f);
    // Restart side file (no counterpart in the reference, whose .pn holds the
    // moments only): the multipliers of every cell, ghosts included, since they
    // are the next closure's warm start.  int64 nx, ny, nm; double time; then
    // alpha[(nx+4)*(ny+4)][nm].
    std::vector<double> alpha(size_t(g.NX) * g.NY * nm);
    if (mnc_grid_get(ctx, nullptr, alpha.data()) != MNC_OK) {
        std::printf("mnc_grid_get: %s\n", mnc_last_error(ctx));
        std::exit(1);
    }
    std::snprintf(name, sizeof name, "out_%.3f_%d.alpha", time, 0);
    f = std::fopen(name, "wb");
    if (!f) { std::printf("cannot write %s\n", name); std::exit(1); }
    int64_t hdr[3] = {g.nx, g.ny, nm};
    std::fwrite(hdr, sizeof(int64_t), 3, f);
    std::fwrite(&time, sizeof(double), 1, f);
    std::fwrite(alpha.data(), sizeof(double), alpha.size(), f);
    if (opt_type != MNC_OPT_NEWTON) {
        // the change-of-basis optimisers continue from each cell's basis and
        // multipliers in that basis (c_P2M, c_alphaP): appended after alpha
        std::vector<double> B(size_t(g.NX) * g.NY * nm * nm), aP(size_t(g.NX) * g.NY * nm);
        if (mnc_grid_get_basis(ctx, B.data(), aP.data()) != MNC_OK) {
            std::printf("mnc_grid_get_basis: %s\n", mnc_last_error(ctx));
            std::exit(1);
        }
        std::fwrite(B.data(), sizeof(double), B.size(), f);
        std::fwrite(aP.data(), sizeof(double), aP.size(), f);
    }
    std::fclose(f);
}

// Reads a restart: <prefix>.pn (the moments of the interior cells) and
// <prefix>.alpha (time and the multipliers of every cell).  Returns false with
// a message if the files do not match the deck's grid.
bool read_restart(const std::string &prefix, const Grid &g, int nm, std::vector<double> &mom,
                  std::vector<double> &alpha, double &time, int opt_type, std::vector<double> &basis,
                  std::vector<double> &alpha_p) {
    FILE *f = std::fopen((prefix + ".pn").c_str(), "rb");
    if (!f) { std::printf("cannot read %s.pn\n", prefix.c_str()); return false; }
    int64_t nx = 0, ny = 0, n = 0;
    double box[4];
    bool ok = std::fread(&nx, sizeof nx, 1, f) == 1 && std::fread(&ny, sizeof ny, 1, f) == 1 &&
              std::fread(box, sizeof(double), 4, f) == 4 && std::fread(&n, sizeof n, 1, f) == 1;
    ok = ok && nx == g.nx && ny == g.ny && n == nm;
    for (int i = 2; ok && i < 2 + g.nx; ++i)
        for (int j = 2; ok && j < 2 + g.ny; ++j)
            ok = std::fread(&mom[(size_t(i) * g.NY + j) * nm], sizeof(double), nm, f) == size_t(nm);
    std::fclose(f);
    if (!ok) { std::printf("%s.pn does not match the deck's grid\n", prefix.c_str()); return false; }
    f = std::fopen((prefix + ".alpha").c_str(), "rb");
    if (!f) { std::printf("cannot read %s.alpha\n", prefix.c_str()); return false; }
    int64_t hdr[3] = {0, 0, 0};
    ok = std::fread(hdr, sizeof(int64_t), 3, f) == 3 && std::fread(&time, sizeof(double), 1, f) == 1;
    ok = ok && hdr[0] == g.nx && hdr[1] == g.ny && hdr[2] == nm;
    ok = ok && std::fread(alpha.data(), sizeof(double), alpha.size(), f) == alpha.size();
    if (opt_type != MNC_OPT_NEWTON) {     // a change-of-basis run's file carries c_P2M and c_alphaP too
        basis.assign(size_t(g.NX) * g.NY * nm * nm, 0.0);
        alpha_p.assign(size_t(g.NX) * g.NY * nm, 0.0);
        ok = ok && std::fread(basis.data(), sizeof(double), basis.size(), f) == basis.size() &&
             std::fread(alpha_p.data(), sizeof(double), alpha_p.size(), f) == alpha_p.size();
    }
    std::fclose(f);
    if (!ok) { std::printf("%s.alpha does not match the deck's grid\n", prefix.c_str()); return false; }
    return true;
}

}  // namespace

int main(int argc, char **argv) {
    mnc_host::InputDeck deck;
    const std::string path = argc > 1 ? argv[1] : "input.deck";
    deck.read(path);
    std::string solver;
    int ncx, ncy, npx, npy, init_cond;
    double ax, bx, ay, by, t_final, out_dt, gsig, floor_v, sigma;
    if (!deck.get("SOLVER", solver)) fail("main.cpp", 69);
    if (!deck.get("NUM_CELLS_X", ncx)) fail("main.cpp", 70);
    if (!deck.get("NUM_CELLS_Y", ncy)) fail("main.cpp", 71);
    if (!deck.get("NUM_MPI_PARTITIONS_X", npx)) fail("main.cpp", 72);
    if (!deck.get("NUM_MPI_PARTITIONS_Y", npy)) fail("main.cpp", 73);
    if (!deck.get("A_X", ax)) fail("main.cpp", 74);
    if (!deck.get("B_X", bx)) fail("main.cpp", 75);
    if (!deck.get("A_Y", ay)) fail("main.cpp", 76);
    if (!deck.get("B_Y", by)) fail("main.cpp", 77);
    if (!deck.get("T_FINAL", t_final)) fail("main.cpp", 78);
    if (!deck.get("OUT_DELTA_T", out_dt)) fail("main.cpp", 79);
    if (!deck.get("GAUSSIAN_SIGMA", gsig)) fail("main.cpp", 80);
    if (!deck.get("FLOOR", floor_v)) fail("main.cpp", 81);
    if (!deck.get("INIT_COND", init_cond)) fail("main.cpp", 82);
    if (!deck.get("SIGMA", sigma)) fail("main.cpp", 83);
    deck.print();
    if (solver != "momopt") {
        std::printf("Solver type does not exist\n.");
        return 1;
    }
    int order, qorder, max_iter, use_gaunt, mtype, otype;
    double cfl, tol, cond_h_max, theta, delta_ppn;
    if (!deck.get("MOMOPT_MOMENT_ORDER", order)) fail("momopt_init.cpp", 56);
    if (!deck.get("MOMOPT_QUAD_ORDER", qorder)) fail("momopt_init.cpp", 57);
    if (!deck.get("MOMOPT_CFL_FACTOR", cfl)) fail("momopt_init.cpp", 58);
    if (!deck.get("MOMOPT_TOL", tol)) fail("momopt_init.cpp", 59);
    if (!deck.get("MOMOPT_COND_H_MAX", cond_h_max)) fail("momopt_init.cpp", 60);
    if (!deck.get("MOMOPT_MAX_ITER", max_iter)) fail("momopt_init.cpp", 62);
    if (!deck.get("MOMOPT_USE_GAUNT", use_gaunt)) fail("momopt_init.cpp", 64);
    if (!deck.get("MOMOPT_THETA", theta)) fail("momopt_init.cpp", 65);
    if (!deck.get("MOMOPT_DELTA_PPN", delta_ppn)) fail("momopt_init.cpp", 66);
    if (!deck.get("MOMOPT_MOMENT_TYPE", mtype)) fail("momopt_init.cpp", 67);
    if (!deck.get("MOMOPT_OPTIMIZATION_TYPE", otype)) fail("momopt_init.cpp", 68);
    if ((mtype != 0 && mtype != 1) || otype < 0 || otype > 2 || (otype != 0 && (mtype != 0 || use_gaunt != 0))) {
        std::printf("built: MOMOPT_MOMENT_TYPE 0 (Mn) or 1 (PPn) with MOMOPT_OPTIMIZATION_TYPE 0, and "
                    "MOMOPT_OPTIMIZATION_TYPE 1 or 2 for Mn with MOMOPT_USE_GAUNT 0\n");
        return 1;
    }
    int max_bfgs_iter = 0;
    double cond_h_max_bfgs = 0.0;
    if (!deck.get("MOMOPT_COND_H_MAX_BFGS", cond_h_max_bfgs)) fail("momopt_init.cpp", 61);
    if (!deck.get("MOMOPT_MAX_BFGS_ITER", max_bfgs_iter)) fail("momopt_init.cpp", 63);
    int prec = MNC_PREC_FP64, warm = MNC_WARM_PREVIOUS, device = 0, ls = MNC_LS_REFERENCE;
    deck.get("MOMOPT_GPU_PRECISION", prec);
    deck.get("MOMOPT_GPU_WARM_START", warm);
    deck.get("MOMOPT_GPU_LINE_SEARCH", ls);
    deck.get("MOMOPT_GPU_DEVICE", device);
    double tol_gamma = 0.0;
    deck.get("MOMOPT_TOL_GAMMA", tol_gamma);

    Grid g;
    g.nx = ncx;
    g.ny = ncy;
    g.ax = ax; g.bx = bx; g.ay = ay; g.by = by;
    g.dx = (bx - ax) / ncx;                          // main.cpp:140-141
    g.dy = (by - ay) / ncy;
    g.init_x = ax;
    g.init_y = ay;
    initialize_grid(g, init_cond, gsig, floor_v, sigma);

    mnc_setup setup{order, qorder, use_gaunt, mtype, device};
    mnc_ctx *ctx = nullptr;
    int st = mnc_create(&setup, &ctx);
    if (st != MNC_OK) { std::printf("mnc_create failed: %d\n", st); return 1; }
    const int nm = mnc_num_moments(ctx);
    const size_t nc = size_t(g.NX) * g.NY;
    std::vector<double> mom(nc * nm, 0.0), alpha(nc * nm, 0.0);
    for (size_t c = 0; c < nc; ++c) {                // momopt_init.cpp:120-143
        mom[c * nm] = g.init[c] * 2.0 * std::sqrt(M_PI);
        if (mtype == 1)                              // PPn (:138-141)
            alpha[c * nm] = mom[c * nm] - 4.0 * delta_ppn * M_PI / mom[c * nm];
        else
            alpha[c * nm] = 2.0 * std::sqrt(M_PI) * std::log(mom[c * nm] / (2.0 * std::sqrt(M_PI)));
    }
    // MOMOPT_GPU_RESTART <prefix>: start from <prefix>.pn and <prefix>.alpha
    // (written at every output time) instead of the initial condition
    std::string restart;
    double t0 = 0.0;
    const bool restarting = deck.get("MOMOPT_GPU_RESTART", restart);
    std::vector<double> basis, alpha_p;
    if (restarting && !read_restart(restart, g, nm, mom, alpha, t0, otype, basis, alpha_p)) return 1;
    mnc_grid_desc gd{g.nx, g.ny, g.dx, g.dy, g.init_x, g.init_y, init_cond, theta, delta_ppn};
    st = mnc_grid_create(ctx, &gd, mom.data(), alpha.data(), g.sigma_t.data(), g.sigma_s.data());
    if (st != MNC_OK) { std::printf("mnc_grid_create: %s\n", mnc_last_error(ctx)); return 1; }
    if (restarting && (st = mnc_ghost_fill(ctx)) != MNC_OK) {        // the ghost moments of the .pn's interior
        std::printf("mnc_ghost_fill: %s\n", mnc_last_error(ctx));
        return 1;
    }
    if (restarting && otype != MNC_OPT_NEWTON &&
        (st = mnc_grid_set_basis(ctx, basis.data(), alpha_p.data())) != MNC_OK) {
        std::printf("mnc_grid_set_basis: %s\n", mnc_last_error(ctx));
        return 1;
    }
    const double max_dt = cfl * (2.0 / (2.0 + theta)) * (g.dx * g.dy) / (g.dx + g.dy);   // :72
    std::printf("--- Init done ---\n");
    mnc_options opt{max_iter, tol, cond_h_max, prec, warm, ls, delta_ppn, otype, max_bfgs_iter, cond_h_max_bfgs,
                    tol_gamma};

    double t = t0, dt = max_dt;                       // main.cpp:204-230
    if (!restarting) write_output(g, ctx, nm, t, otype);
    for (double t_out = std::fmin(t + out_dt, t_final); t_out <= t_final; t_out = std::fmin(t_out + out_dt, t_final)) {
        dt = max_dt;
        while (t < t_out) {
            t += dt;
            if (t > t_out) {
                dt = t_out - (t - dt);
                t = t_out;
            }
            st = mnc_update(ctx, dt, &opt);
            if (st != MNC_OK) {
                std::printf("%s\n", mnc_last_error(ctx));
                return 1;                             // the reference aborts here
            }
            if ((t * 100 - std::floor(t * 100)) / 100.0 < dt) std::printf("dt = %f   t = %f\n", dt, t);
        }
        write_output(g, ctx, nm, t, otype);
        if (t_out > t_final - 1e-12 * t_final) break;
    }
    mnc_destroy(ctx);
    return 0;
}
