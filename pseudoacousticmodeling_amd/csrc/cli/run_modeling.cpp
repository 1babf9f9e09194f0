// run_modeling -- forward modelling of every shot of a survey on the GPU, the
// drop-in for the C++ reference's run_modeling (cpp/modeling_main.cpp:5-20,
// cpp/Modeling.cu:433-479).  Usage: run_modeling [Parameters.json]
// (default ../inputs/Parameters.json, the reference's fixed path);
// `run_modeling <Parameters.json> --check` prints the parsed survey only.
// Writes seismogram_shot_{s+1}_Nt{nt_data}_Nrec{Nrec}_fcut{fcut:.1f}.bin,
// raw float32 [nt_data][Nrec], into seismogramFolder.  With "snap": true it
// also writes the interior [nz][nx] of the field after step k for every k
// that is a multiple of "step" up to "last_save", as
// {approx}forward_shot_{s+1}Nx{nx}_Nz{nz}_Nt{nt}_frame{k}.bin in
// snapshotFolder; as in cpp/Modeling.cu:433-470 the last one taken is
// written under the name frame{last_save}.
#include <chrono>
#include <iostream>

#include "pam_cli.h"

// Snapshots through the per-step interface: one shot at a time in slot 0.
static void write_snapshots(const cli::Survey &s, pam_ctx *ctx) {
    if (s.step <= 0) throw std::invalid_argument("snap needs step > 0");
    const size_t np = (size_t)s.nz_abc * s.nx_abc;
    float *d_field = nullptr;
    CLI_HIP(hipMalloc(&d_field, np * sizeof(float)));
    std::vector<float> pad(np), frame((size_t)s.nz * s.nx);
    auto name = [&](int shot, int k) {
        return s.snapshotFolder + s.approximation + "forward_shot_" + std::to_string(shot + 1) + "Nx" +
               std::to_string(s.nx) + "_Nz" + std::to_string(s.nz) + "_Nt" + std::to_string(s.nt) + "_frame" +
               std::to_string(k) + ".bin";
    };
    for (int shot = 0; shot < s.Nshot; ++shot) {
        CLI_PAM(ctx, pam_reset(ctx, shot));
        int held = -1;  // the frame taken last and not yet written
        for (int k = 0; k < s.nt; ++k) {
            CLI_PAM(ctx, pam_forward_step(ctx, k, nullptr));
            if (k <= s.last_save && k % s.step == 0) {
                if (held >= 0) cli::write_bin(name(shot, held), frame.data(), frame.size());
                CLI_PAM(ctx, pam_get_field(ctx, 1, d_field));
                CLI_HIP(hipMemcpy(pad.data(), d_field, np * sizeof(float), hipMemcpyDeviceToHost));
                for (int z = 0; z < s.nz; ++z)
                    std::copy(pad.begin() + (size_t)(z + s.N_abc) * s.nx_abc + s.N_abc,
                              pad.begin() + (size_t)(z + s.N_abc) * s.nx_abc + s.N_abc + s.nx,
                              frame.begin() + (size_t)z * s.nx);
                held = k;
            }
        }
        if (held >= 0) cli::write_bin(name(shot, s.last_save), frame.data(), frame.size());
    }
    (void)hipFree(d_field);
}

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
        const bool stability_only = argc > 2 && std::string(argv[2]) == "--stability";
        if (!stability_only)
            std::cout << "info: Solving " << s.approximation << " wave equation, " << s.Nshot << " shots, " << s.nx
                      << "x" << s.nz << ", nt=" << s.nt << std::endl;
        const size_t nm = (size_t)s.nx * s.nz;
        std::vector<float> vp = cli::read_bin(s.vpFile, nm), eps, del, th;
        if (s.approximation != "acoustic") {
            eps = cli::read_bin(s.epsilonFile, nm);
            del = cli::read_bin(s.deltaFile, nm);
        }
        if (s.approximation == "TTI") {
            th = cli::read_bin(s.thetaFile, nm);
            for (float &v : th) v = v * cli::kPi / 180.0f;  // degrees on disk (cpp/Modeling.cu:364-367)
        }
        // the reference checks as it sets the model (cpp/Modeling.cu:371); --stability: the check alone
        cli::check_dispersion_stability(s, vp, eps, del, th);
        if (stability_only) return 0;
        pam_ctx *ctx = cli::make_context(s);
        float *d_vp = cli::to_device(vp);
        float *d_eps = eps.empty() ? nullptr : cli::to_device(eps);
        float *d_del = del.empty() ? nullptr : cli::to_device(del);
        float *d_th = th.empty() ? nullptr : cli::to_device(th);
        CLI_PAM(ctx, pam_set_model(ctx, d_vp, d_eps, d_del, d_th, PAM_INTERIOR));
        const size_t nd = (size_t)s.nt_data * s.Nrec;
        float *d_seis = nullptr;
        CLI_HIP(hipMalloc(&d_seis, nd * s.Nshot * sizeof(float)));
        CLI_PAM(ctx, pam_forward(ctx, 0, s.Nshot, d_seis));
        std::vector<float> h(nd);
        for (int shot = 0; shot < s.Nshot; ++shot) {
            CLI_HIP(hipMemcpy(h.data(), d_seis + shot * nd, nd * sizeof(float), hipMemcpyDeviceToHost));
            cli::write_bin(s.seismogramFolder + "seismogram_shot_" + std::to_string(shot + 1) + "_Nt" +
                               std::to_string(s.nt_data) + "_Nrec" + std::to_string(s.Nrec) + "_fcut" + s.fcut_text() +
                               ".bin",
                           h.data(), nd);
        }
        if (s.snap) write_snapshots(s, ctx);
        double loop_ms = 0, kern_ms = 0;
        long long launches = 0;
        pam_last_timing(ctx, &loop_ms, &kern_ms, &launches);
        std::cout << "info: GPU time stepping " << loop_ms << " ms" << std::endl;
        for (float *p : {d_vp, d_eps, d_del, d_th, d_seis})
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
