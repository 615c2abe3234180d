// The reference's driver loop (test.cpp:9-152) written against the drop-in
// classes of include/splash/MUSCL.hpp: the same input files, writers, four
// light-curve files and stop logic.  Two departures: the parameter file is
// read by a flat-key reader instead of nlohmann::json (the same keys and
// values), and the mesh is chosen by arguments <kind> <level> [nx ny]
// instead of being fixed to quad_sphere(6) in the source.
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "splash/MUSCL.hpp"

using namespace splash;

static double key(const std::string& text, const std::string& k) {
    const size_t p = text.find("\"" + k + "\"");
    if (p == std::string::npos) throw std::runtime_error("missing key " + k);
    const size_t c = text.find(':', p);
    std::string v = text.substr(c + 1, text.find_first_of(",}\n", c + 1) - c - 1);
    if (v.find("true") != std::string::npos) return 1;
    if (v.find("false") != std::string::npos) return 0;
    return std::stod(v);
}

int main(int argc, char** argv) {
    const std::string kind = argc > 1 ? argv[1] : "quad";
    const int level = argc > 2 ? std::stoi(argv[2]) : 6;
    SurfaceMesh mesh = kind == "hex" ? icosphere_hex(size_t(level))
                     : kind == "ico" ? icosphere(size_t(level))
                     : kind == "uv"  ? uv_sphere(level, argc > 3 ? std::stoi(argv[3]) : level / 2)
                                     : quad_sphere(size_t(level));

    std::ifstream ifs("input/parameters.json");
    std::string text((std::istreambuf_iterator<char>(ifs)), std::istreambuf_iterator<char>());
    const double dt = key(text, "dt"), t_max = key(text, "t_max");
    size_t maxstep = size_t(key(text, "maxstep")), skipstep = size_t(key(text, "skipstep"));
    const size_t dim = size_t(key(text, "dim"));
    const double gam3d = key(text, "gam3d"), gam = 2 - 1 / gam3d, omega_ns = key(text, "omega_ns");
    const size_t threads = size_t(key(text, "threads"));
    const bool accretion_on = key(text, "accretion_on") != 0;

    // test.cpp:34-77: read input/input.dat, pad with 1 and do no steps if short
    std::ifstream inData("input/input.dat");
    std::vector<double> temp;
    double element;
    while (inData >> element) temp.push_back(element);
    if (temp.size() < mesh.n_faces() * dim) {
        temp.resize(mesh.n_faces() * dim, 1.0);
        maxstep = 0;
        std::cout << "input file does not have enough values!" << std::endl;
    }
    std::vector<std::vector<double>> U_in(mesh.n_faces(), std::vector<double>(dim));
    for (size_t i = 0; i < mesh.n_faces(); i++)
        for (size_t j = 0; j < dim; j++) U_in[i][j] = temp[i * dim + j];

    MUSCL_HLLCplus test2(mesh, U_in, int(dim), gam, omega_ns, accretion_on, threads);
    test2.write_face_centers();
    test2.write_faces();
    test2.write_vertices();
    test2.write_t_rho();
    test2.write_t_p();
    test2.write_t_omega_z();
    test2.write_t_L();

    std::ofstream out_lc_0("results/lightcurve0.dat");
    std::ofstream out_lc_45("results/lightcurve45.dat");
    std::ofstream out_lc_90("results/lightcurve90.dat");
    std::ofstream out_lc_180("results/lightcurve180.dat");
    auto write_light_curves = [&]() {
        const std::vector<double> lc = test2.get_light_curves();
        out_lc_0 << test2.time() << " " << lc[0] << "\n";
        out_lc_0.flush();
        out_lc_45 << test2.time() << " " << lc[1] << "\n";
        out_lc_45.flush();
        out_lc_90 << test2.time() << " " << lc[2] << "\n";
        out_lc_90.flush();
        out_lc_180 << test2.time() << " " << lc[3] << "\n";
        out_lc_180.flush();
    };
    write_light_curves();
    size_t steps = 0;
    while (test2.time() < t_max && steps < maxstep) {
        test2.do_step(dt);
        write_light_curves();
        if (steps % skipstep == 0) {
            test2.write_t_rho();
            test2.write_t_p();
            test2.write_t_curl();
            test2.write_t_omega_z();
            test2.write_t_L();
        }
        if (test2.get_stop_check()) {
            test2.write_t_rho();
            test2.write_t_p();
            test2.write_t_curl();
            test2.write_t_omega_z();
            test2.write_t_L();
            break;
        }
        steps++;
    }
    test2.write_final_state();
    return 0;
}
