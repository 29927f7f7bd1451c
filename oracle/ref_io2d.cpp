// TEST INFRASTRUCTURE ONLY — the reference's own 2D Sod program
// (examples/sod2d/sod2d.cpp), compiled from its source with `main` renamed,
// driven to call its .dat writers (write_output_2d and write_xslice,
// sod2d.cpp:58-152) on a given state. Used to produce the golden files that
// pin nest_amd's writers.
#define main ref_sod2d_program_main
#include "sod2d.cpp"
#undef main

#include <unistd.h>

extern "C" int ref_write_sod2d(const double* G, int nx, int ny, int npx, int npy, double x_min, double x_max,
                               double y_min, double y_max, double gamma, double time, int step, const char* dir) {
    config_t cfg;
    cfg.x_min = x_min;
    cfg.x_max = x_max;
    cfg.y_min = y_min;
    cfg.y_max = y_max;
    cfg.nx_zones = nx;
    cfg.ny_zones = ny;
    cfg.nx_patches = npx;
    cfg.ny_patches = npy;
    cfg.gamma = gamma;
    auto domain = index_space(ivec(0, 0), uvec(static_cast<unsigned int>(nx), static_cast<unsigned int>(ny)));
    auto patches = build_periodic_mesh_2d<double>(domain, npx, npy, 4, 2);
    for (auto& patch : patches)
        for (auto idx : patch.interior)
            for (std::size_t k = 0; k < 4; ++k)
                patch(k, idx) = G[(k * (std::size_t)nx + idx[0]) * (std::size_t)ny + idx[1]];
    char cwd[4096];
    if (!getcwd(cwd, sizeof cwd) || chdir(dir) != 0) return 1;
    write_output_2d(patches, cfg, time, step);
    write_xslice(patches, cfg, time, step);
    return chdir(cwd) == 0 ? 0 : 1;
}
