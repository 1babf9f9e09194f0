// run_migration -- reverse-time migration of every shot on the GPU, the
// drop-in for the C++ reference's run_migration (cpp/migration_main.cpp:6-33,
// cpp/Migration.cu:359-397): smoothed model (cpp/Migration.cu:252-310), muted
// data (cpp/Migration.cu:86-128), adjoint steps t = nt-1 .. 0, image / ilum
// where ilum > 1e-12 (cpp/Migration.cu:470-484).  "migration": "onthefly"
// keeps every source-field step in HBM, "checkpoint" keeps uniform
// checkpoints every "step" steps in HBM, "SB" keeps the interior rims of every
// step and rebuilds the source field backwards.  Smoothing and mute run on the
// GPU.  Usage: run_migration [Parameters.json]
// (`--check` after the path prints the parsed survey only).
// Writes image_{approx}_Nx{nx}_Nz{nz}.bin, raw float32 [nz][nx], into imageFolder.
#include <chrono>
#include <iostream>

#include "pam_cli.h"

int main(int argc, char **argv) {
    const auto t0 = std::chrono::system_clock::now();
    try {
        const bool check = argc > 2 && std::string(argv[2]) == "--check";
        cli::Survey s(argc > 1 ? argv[1] : "../inputs/Parameters.json");
        if (check) {  // the parsed survey, no GPU work
            std::cout << "nx " << s.nx << " nz " << s.nz << " nt " << s.nt << " itlag " << s.itlag << " nt_data "
                      << s.nt_data << " Nshot " << s.Nshot << " Nrec " << s.Nrec << " fcut " << s.fcut_text() << "\n";
            for (int i = 0; i < s.Nshot; ++i) std::cout << "src " << s.sz[i] << " " << s.sx[i] << "\n";
            for (int i = 0; i < s.Nrec; ++i) std::cout << "rec " << s.rz[i] << " " << s.rx[i] << "\n";
            return 0;
        }
        std::cout << "info: Solving " << s.approximation << " Reverse Time Migration by " << s.migration
                  << " method." << std::endl;
        int storage = PAM_STORE_ALL_F32;
        if (s.migration == "checkpoint") storage = PAM_STORE_CHECKPOINT;
        else if (s.migration == "SB") storage = PAM_STORE_BOUNDARY_F32;
        else if (s.migration != "onthefly") {
            std::cerr << "Error: Invalid migration method: " << s.migration << std::endl;
            return 1;
        }
        const size_t nm = (size_t)s.nx * s.nz;
        std::vector<float> vp = cli::read_bin(s.vpFile, nm), eps, del, th;
        const std::vector<char> mask = cli::water_mask(vp);
        cli::smooth(s, vp, mask, false);
        if (s.approximation != "acoustic") {
            eps = cli::read_bin(s.epsilonFile, nm);
            del = cli::read_bin(s.deltaFile, nm);
            cli::smooth(s, eps, mask, true);
            cli::smooth(s, del, mask, true);
        }
        if (s.approximation == "TTI") {
            th = cli::read_bin(s.thetaFile, nm);
            cli::smooth(s, th, mask, true);
            for (float &v : th) v = v * cli::kPi / 180.0f;
        }
        pam_ctx *ctx = cli::make_context(s);
        CLI_PAM(ctx, pam_set_option(ctx, PAM_OPT_STORAGE, storage));
        if (storage == PAM_STORE_CHECKPOINT) CLI_PAM(ctx, pam_set_option(ctx, PAM_OPT_CKPT_STEP, s.step));
        CLI_PAM(ctx, pam_set_option(ctx, PAM_OPT_ADJ_RANGE, 1));
        CLI_PAM(ctx, pam_set_option(ctx, PAM_OPT_NORMALIZE, 2));
        float *d_vp = cli::to_device(vp);
        float *d_eps = eps.empty() ? nullptr : cli::to_device(eps);
        float *d_del = del.empty() ? nullptr : cli::to_device(del);
        float *d_th = th.empty() ? nullptr : cli::to_device(th);
        CLI_PAM(ctx, pam_set_model(ctx, d_vp, d_eps, d_del, d_th, PAM_INTERIOR));
        const size_t nd = (size_t)s.nt_data * s.Nrec;
        std::vector<float> data(nd * s.Nshot);
        for (int shot = 0; shot < s.Nshot; ++shot) {
            std::vector<float> g = cli::read_bin(s.seismogramFolder + "seismogram_shot_" + std::to_string(shot + 1) +
                                                     "_Nt" + std::to_string(s.nt_data) + "_Nrec" +
                                                     std::to_string(s.Nrec) + "_fcut" + s.fcut_text() + ".bin",
                                                 nd);
            std::copy(g.begin(), g.end(), data.begin() + shot * nd);
        }
        float *d_data = cli::to_device(data), *d_img = nullptr, *d_ilum = nullptr;
        cli::mute_all(s, d_data);
        CLI_HIP(hipMalloc(&d_img, nm * sizeof(float)));
        CLI_HIP(hipMalloc(&d_ilum, nm * sizeof(float)));
        CLI_PAM(ctx, pam_rtm(ctx, 0, s.Nshot, d_data, d_img, d_ilum));
        std::vector<float> img(nm);
        CLI_HIP(hipMemcpy(img.data(), d_img, nm * sizeof(float), hipMemcpyDeviceToHost));
        cli::write_bin(s.imageFolder + "image_" + s.approximation + "_Nx" + std::to_string(s.nx) + "_Nz" +
                           std::to_string(s.nz) + ".bin",
                       img.data(), nm);
        for (float *p : {d_vp, d_eps, d_del, d_th, d_data, d_img, d_ilum})
            if (p) (void)hipFree(p);
        pam_destroy(ctx);
    } catch (const std::exception &e) {
        std::cerr << "Error: " << e.what() << std::endl;
        return 1;
    }
    const std::chrono::duration<double> el = std::chrono::system_clock::now() - t0;
    std::cout << "\nRun time: " << el.count() << " s." << std::endl;
    return 0;
}
