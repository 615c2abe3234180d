// Prints the physics keys the drop-in adiabatic constructor takes from a
// parameters.json (the keys of adiabatic.hpp:59-68), parsed by the drop-in
// reader of splash/MUSCL.hpp, one "key value" line each.
#include <cstdio>
#include <exception>

#include "splash/MUSCL.hpp"

int main(int argc, char** argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: params_reader <parameters.json>\n");
        return 2;
    }
    try {
        const splash_params p = splash::read_parameters_json(argv[1]);
        std::printf("accretion_rate %.17g\n", p.acc_rate);
        std::printf("e_acc %.17g\n", p.e_acc);
        std::printf("omega_acc_abs %.17g\n", p.omega_acc_abs);
        std::printf("tilt_angle %.17g\n", p.tilt_angle);
        std::printf("acc_width %.17g\n", p.acc_width);
        std::printf("alpha %.17g\n", p.alpha);
        std::printf("friction_on %d\n", p.friction_on);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
