// amressif_run -- the reference's main program (Source/src1_rever/main.cpp,
// main_main) on the C-ABI of include/amressif.h: read the inputs file,
// start from rest (hybrid_grid_init, main.cpp:793) or from the checkpoint
// chk_out (main.cpp:326-328, 386-800), advance steps starting_step..nsteps
// (main.cpp:875), write a checkpoint every chk_int steps (main.cpp:1167-1170)
// and the flux and flow-field plotfiles every plot_int steps (main.cpp:1045-1048,
// 1185-1191), print the
// per-step report and the run time (main.cpp:1173-1175, 1207-1212).
//
//   amressif_run <inputs> [key=value ...]   (files go to the working directory)
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/amressif.h"

static int die(amressif_ctx *ctx, const char *what, int rc)
{
    std::fprintf(stderr, "amressif_run: %s failed (%d): %s\n", what, rc, ctx ? amressif_last_error(ctx) : "");
    if (ctx) amressif_destroy(ctx);
    return 1;
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <inputs> [key=value ...]\n", argv[0]);
        return 2;
    }
    const auto t_start = std::chrono::steady_clock::now();
    amressif_inputs in;
    char err[512];
    if (amressif_parse_inputs(argv[1], argc - 2, (const char *const *)(argv + 2), &in, err, sizeof err) != AMRESSIF_OK) {
        std::fprintf(stderr, "amressif_run: %s\n", err);
        return 2;
    }
    amressif_ctx *ctx = nullptr;
    int rc = amressif_create(&in, nullptr, &ctx);
    if (rc != AMRESSIF_OK) return die(ctx, "create", rc);
    const double dt = amressif_dt(&in);
    std::printf("PARAMS| n_cell %d, dt %.17g, reynolds number %g, precision %d\n", in.n_cell, dt, in.ren, in.precision);

    double time = 0.0;
    int starting_step = 1;
    if (in.chk_out > 0) {
        char name[64];
        std::snprintf(name, sizeof name, "checkpoint%05d", in.chk_out);
        if ((rc = amressif_read_checkpoint(ctx, name, &time)) != AMRESSIF_OK) return die(ctx, "read_checkpoint", rc);
        std::printf("INFO| restart from %s at time %.17g\n", name, time);
        starting_step = in.chk_out + 1;
    } else {
        if ((rc = amressif_init_state(ctx, nullptr, 0)) != AMRESSIF_OK) return die(ctx, "init_state", rc);
        if (in.plot_int > 0 && (rc = amressif_write_plotfile(ctx, ".", "pltInit", 0, time)) != AMRESSIF_OK)
            return die(ctx, "write_plotfile", rc);
    }

    // advance in chunks that end at the next checkpoint or plotfile step
    std::vector<amressif_step_stats> st;
    for (int n = starting_step; n <= in.nsteps;) {
        int last = in.nsteps;
        for (int every : {in.chk_int, in.plot_int})
            if (every > 0) last = std::min(last, (n + every - 1) / every * every);
        const int k = last - n + 1;
        st.resize(k);
        const auto t0 = std::chrono::steady_clock::now();
        if ((rc = amressif_advance(ctx, k, st.data())) != AMRESSIF_OK) return die(ctx, "advance", rc);
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        for (int s = 0; s < k; ++s) {
            time += dt;
            std::printf("STEP| %d time %.10g countIter %d normError %.6e V-cycles %d |D(u*)| %.6e |D(u)| %.6e "
                        "flags %d\n",
                        n + s, time, st[s].pseudo_iters, st[s].norm_error, st[s].vcycles, st[s].div_star,
                        st[s].div_new, st[s].flags);
        }
        std::printf("INFO| Advancing steps %d..%d took %g seconds\n", n, last, sec);
        if (in.chk_int > 0 && last % in.chk_int == 0 &&
            (rc = amressif_write_checkpoint(ctx, ".", last, time)) != AMRESSIF_OK)
            return die(ctx, "write_checkpoint", rc);
        if (in.plot_int > 0 && last % in.plot_int == 0) {
            // Export_Fluxes at the end of the momentum solve (main.cpp:1045-1048)
            if ((rc = amressif_write_flux_plotfile(ctx, ".", last, time)) != AMRESSIF_OK)
                return die(ctx, "write_flux_plotfile", rc);
            if ((rc = amressif_write_plotfile(ctx, ".", "pltResults", last, time)) != AMRESSIF_OK)
                return die(ctx, "write_plotfile", rc);
        }
        n = last + 1;
    }
    amressif_destroy(ctx);
    std::printf("Run time = %g\n",
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count());
    return 0;
}
