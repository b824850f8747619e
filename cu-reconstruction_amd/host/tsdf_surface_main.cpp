// tsdf_surface_main.cpp -- the reference's TSDF example driver
// (examples/tsdf_surface_reconstruction/tsdf_surface.cpp): parameter file ->
// first RGB-D pair of the dataset directory -> TSDFSurface -> ASCII PLY.
//
//   tsdf_surface [--config_path=FILE]   (default ./data/tsdf_surface_parameters.yaml)
#include "recon/tsdf_surface.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

int main(int argc, char* argv[]) {
    std::string config_path = "./data/tsdf_surface_parameters.yaml";
    for (int i = 1; i < argc; ++i) {
        const char* a = argv[i];
        if (std::strncmp(a, "--config_path=", 14) == 0 || std::strncmp(a, "-config_path=", 13) == 0) {
            config_path = std::strchr(a, '=') + 1;
        } else if ((std::strcmp(a, "--config_path") == 0 || std::strcmp(a, "-config_path") == 0) &&
                   i + 1 < argc) {
            config_path = argv[++i];
        } else {
            std::fprintf(stderr, "usage: %s [--config_path=FILE]\n", argv[0]);
            return EXIT_FAILURE;
        }
    }
    cuphoto::Config cfg;
    if (!cfg.set_parameter_file(config_path)) return EXIT_FAILURE;
    cuphoto::RGBDDataset dataset(cfg.get<std::string>("rgbd.dataset.path"),
                                 (double)cfg.get<float>("depth_scale"));
    // dataset.color.mode: 1 reads the colour image too (an 8-bit RGB PNG of
    // the depth's size), integrates colour and colours the mesh vertices
    const bool color_mode = cfg.get<int>("dataset.color.mode", 0) != 0;
    cuphoto::PtrStepSz<float> depth;
    cuphoto::PtrStepSz<unsigned char> color;
    if (!dataset.get_next(&depth, color_mode ? &color : nullptr)) return EXIT_FAILURE;
    if (color_mode && (color.rows != depth.rows || color.cols != depth.cols)) {
        std::fprintf(stderr, "[cuphoto] colour and depth images differ in size\n");
        return EXIT_FAILURE;
    }
    const std::string store_path = cfg.get<std::string>("mesh_store.path");
    cuphoto::TSDFSurface tsdf(cfg);
    const bool done = color_mode ? tsdf.reconstruct_surface(color, depth) : tsdf.reconstruct_surface(depth);
    if (!done) {
        std::fprintf(stderr, "[cuphoto] reconstruction failed\n");
        return EXIT_FAILURE;
    }
    const cuphoto::Mesh& mesh = tsdf.get_mesh();
    if (!mesh.store_to_ply(store_path)) {
        std::fprintf(stderr, "[cuphoto] cannot write %s\n", store_path.c_str());
        return EXIT_FAILURE;
    }
    std::printf("%zu vertices, %zu triangles -> %s\n", mesh.vertices.size(), mesh.polygons.size(),
                store_path.c_str());
    return EXIT_SUCCESS;
}
