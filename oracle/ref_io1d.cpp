// TEST INFRASTRUCTURE ONLY — the reference's own 1D Sod program
// (examples/sod1d/sod1d.cpp), compiled from its source with `main` renamed,
// driven to call its .dat writer (write_output, sod1d.cpp:50-91) on a given
// state. Used to produce the golden files that pin nest_amd's writers.
#define main ref_sod1d_program_main
#include "sod1d.cpp"
#undef main

#include <unistd.h>

extern "C" int ref_write_sod1d(const double* G, int n, int npatch, double x_min, double x_max, double gamma,
                               double time, int step, const char* dir) {
    config_t cfg;
    cfg.x_min = x_min;
    cfg.x_max = x_max;
    cfg.num_zones = n;
    cfg.num_patches = npatch;
    cfg.gamma = gamma;
    auto domain = index_space(ivec(0), uvec(static_cast<unsigned int>(n)));
    auto patches = build_periodic_mesh_1d<double>(domain, npatch, NCONS, 2);
    for (auto& patch : patches)
        for (auto idx : patch.interior)
            for (std::size_t k = 0; k < NCONS; ++k) patch(k, idx) = G[k * n + idx[0]];
    char cwd[4096];
    if (!getcwd(cwd, sizeof cwd) || chdir(dir) != 0) return 1;
    write_output(patches, cfg, time, step);
    return chdir(cwd) == 0 ? 0 : 1;
}
