// Checkpoint / resume through the drop-in classes: a run that writes a
// checkpoint after n1 steps and continues n2 more must end where a second
// solver that reads the checkpoint and steps n2 ends.  Prints the largest
// difference and exits 0 when the states, times and stop flags are equal.
// Arguments: <precision 0|1> <checkpoint path>.
#include <cmath>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "splash/MUSCL.hpp"

using namespace splash;

int main(int argc, char** argv) {
    const int precision = argc > 1 ? std::stoi(argv[1]) : SPLASH_FP64;
    const std::string path = argc > 2 ? argv[2] : "checkpoint.bin";
    SurfaceMesh mesh = icosphere_hex(4);
    const double gam = 2 - 1 / (4. / 3);
    // solid-body rotation about z with a density bump
    std::vector<std::vector<double>> U(mesh.n_faces(), std::vector<double>(5));
    const auto& x = mesh.vertices();
    const auto& fo = mesh.face_offsets();
    const auto& fv = mesh.face_vertices();
    for (size_t f = 0; f < mesh.n_faces(); ++f) {
        double c[3] = {0, 0, 0};
        for (int32_t k = fo[f]; k < fo[f + 1]; ++k)
            for (int d = 0; d < 3; ++d) c[d] += x[3 * size_t(fv[size_t(k)]) + size_t(d)];
        const double n = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        for (double& v : c) v /= n;
        const double sigma = 1 + 0.1 * std::exp(-20 * ((c[0] - 1) * (c[0] - 1) + c[1] * c[1] + c[2] * c[2]));
        const double om = 0.3, vx = -om * c[1], vy = om * c[0];
        U[f] = {sigma, sigma * (c[1] * 0 - c[2] * vy), sigma * (c[2] * vx - c[0] * 0), sigma * (c[0] * vy - c[1] * vx),
                0.1 * sigma / (gam - 1) + 0.5 * sigma * (vx * vx + vy * vy)};
    }
    splash_params p;
    splash_default_params(&p);
    p.precision = precision;
    MUSCL_HLLCplus a(mesh, U, 5, gam, 0.0, p), b(mesh, U, 5, gam, 0.0, p);
    a.set_verbose(false);
    b.set_verbose(false);
    for (int s = 0; s < 13; ++s) a.do_step(1e-4);
    a.write_checkpoint(path);
    a.do_steps(1e-4, 11);
    b.read_checkpoint(path);
    b.do_steps(1e-4, 11);
    const auto A = a.get_state(), B = b.get_state();
    double diff = 0;
    for (size_t f = 0; f < A.size(); ++f)
        for (size_t k = 0; k < 5; ++k) diff = std::max(diff, std::abs(A[f][k] - B[f][k]));
    std::cout << "max |diff| " << diff << " t " << a.time() << " " << b.time() << "\n";
    bool bad_tag = false;
    try {
        b.read_checkpoint(path + ".missing");
    } catch (const std::runtime_error&) {
        bad_tag = true;
    }
    // the next diagnostics line (time, mass gain against the first step's
    // mass, angular momentum) is the same on the resumed run
    std::string line[2];
    for (int k = 0; k < 2; ++k) {
        MUSCL_HLLCplus& s = k ? b : a;
        std::ostringstream capture;
        std::streambuf* old = std::cout.rdbuf(capture.rdbuf());
        s.set_verbose(true);
        s.do_step(1e-4);
        std::cout.rdbuf(old);
        line[k] = capture.str();
    }
    std::cout << line[0] << line[1];
    const bool same_line = line[0] == line[1] && line[0].find("inf") == std::string::npos;
    // a mesh with the same face count but one vertex moved cannot read it
    std::vector<double> xyz = mesh.vertices();
    xyz[0] += 1e-12;
    SurfaceMesh moved(xyz, mesh.face_offsets(), mesh.face_vertices());
    MUSCL_HLLCplus c(moved, U, 5, gam, 0.0, p);
    bool other_mesh = false;
    try {
        c.read_checkpoint(path);
    } catch (const std::runtime_error&) {
        other_mesh = true;
    }
    std::cout << "same line " << same_line << " bad tag " << bad_tag << " other mesh rejected " << other_mesh << "\n";
    return diff == 0 && a.time() == b.time() && a.get_stop_check() == b.get_stop_check() && bad_tag && same_line &&
                   other_mesh
               ? 0
               : 1;
}
