// The C++ ensemble wrapper against single SPLASH_MIXED drop-in solvers:
// three runs on different meshes and rotations, 20 steps each; exits 0 when
// every run's state and time equal its single solver's bit for bit.
#include <cmath>
#include <iostream>
#include <vector>

#include "splash/Ensemble.hpp"

using namespace splash;

static std::vector<std::vector<double>> rotating(const SurfaceMesh& mesh, double om, double gam) {
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
        const double sigma = 1 + 0.05 * c[0], vx = -om * c[1], vy = om * c[0];
        U[f] = {sigma, -sigma * c[2] * vy, sigma * c[2] * vx, sigma * (c[0] * vy - c[1] * vx),
                0.1 * sigma / (gam - 1) + 0.5 * sigma * (vx * vx + vy * vy)};
    }
    return U;
}

int main() {
    const double gam = 2 - 1 / (4. / 3);
    splash_params p;
    splash_default_params(&p);
    p.precision = SPLASH_MIXED;
    std::vector<MUSCL_ensemble::Run> runs = {{icosphere_hex(3), {}, gam, 0.0, p},
                                       {quad_sphere(4), {}, gam, 0.1, p},
                                       {icosphere(3), {}, gam, 0.0, p}};
    const double om[3] = {0.2, 0.35, 0.1};
    for (size_t r = 0; r < runs.size(); ++r) runs[r].U_in = rotating(runs[r].mesh, om[r], gam);
    MUSCL_ensemble ens(runs, 5, SPLASH_HLLCPLUS);
    const std::vector<double> dt0 = {1e-4, 2e-4, 1e-4};
    ens.do_steps(dt0, 20);
    int bad = 0;
    for (size_t r = 0; r < runs.size(); ++r) {
        MUSCL_HLLCplus single(runs[r].mesh, runs[r].U_in, 5, gam, runs[r].omega_ns, p);
        single.set_verbose(false);
        single.do_steps(dt0[r], 20);
        const auto A = single.get_state(), B = ens.get_state(int(r));
        double d = 0;
        for (size_t f = 0; f < A.size(); ++f)
            for (size_t k = 0; k < 5; ++k) d = std::max(d, std::abs(A[f][k] - B[f][k]));
        std::cout << "run " << r << " max |diff| " << d << " t " << ens.time(int(r)) << "\n";
        bad += d != 0 || single.time() != ens.time(int(r));
    }
    return bad;
}
