// C++ wrapper of the batched-ensemble C ABI (splash_ensemble_* in
// include/splash.h).  No counterpart in the reference, which advances one
// run per solver object: here many independent runs of one solver and dim,
// each with its own mesh, state, gam, omega_ns and parameters, advance with
// one launch per RK stage for all of them (FP32 flux / FP64 state), and
// each run evolves exactly as a single SPLASH_MIXED solver would.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "../splash.h"
#include "MUSCL.hpp"

namespace splash {

class MUSCL_ensemble {
public:
    struct Run {
        SurfaceMesh mesh;
        std::vector<std::vector<double>> U_in;  // [n_faces][dim]
        double gam;
        double omega_ns;
        splash_params params;
    };

    // solver: SPLASH_HLLE (dim 4) or SPLASH_HLLE_P / SPLASH_HLLC / SPLASH_HLLCPLUS (dim 5)
    MUSCL_ensemble(const std::vector<Run>& runs, int dim, int solver) : dim_(dim) {
        const size_t n = runs.size();
        std::vector<splash_mesh_desc> desc(n);
        std::vector<std::vector<double>> flat(n);
        std::vector<const double*> U(n);
        std::vector<double> gam(n), om(n);
        std::vector<splash_params> p(n);
        for (size_t r = 0; r < n; ++r) {
            desc[r] = runs[r].mesh.desc();
            if (runs[r].U_in.size() != runs[r].mesh.n_faces())
                throw std::invalid_argument("U_in of run " + std::to_string(r) + " is the wrong size");
            flat[r].assign(runs[r].mesh.n_faces() * size_t(dim), 0.0);
            for (size_t i = 0; i < runs[r].U_in.size(); ++i)
                for (int k = 0; k < dim && size_t(k) < runs[r].U_in[i].size(); ++k)
                    flat[r][i * size_t(dim) + size_t(k)] = runs[r].U_in[i][size_t(k)];
            U[r] = flat[r].data();
            gam[r] = runs[r].gam;
            om[r] = runs[r].omega_ns;
            p[r] = runs[r].params;
            faces_.push_back(runs[r].mesh.n_faces());
        }
        check_rc(splash_ensemble_create(int(n), desc.data(), U.data(), dim, solver, gam.data(), om.data(), p.data(),
                                        &e_));
    }
    ~MUSCL_ensemble() { splash_ensemble_destroy(e_); }
    MUSCL_ensemble(const MUSCL_ensemble&) = delete;
    MUSCL_ensemble& operator=(const MUSCL_ensemble&) = delete;

    int n_runs() const { return int(faces_.size()); }
    // nsteps do_step calls of every run; dt0[r] is run r's requested step
    void do_steps(const std::vector<double>& dt0, int nsteps) {
        if (int(dt0.size()) != n_runs()) throw std::invalid_argument("one dt0 per run");
        check_rc(splash_ensemble_step(e_, dt0.data(), nsteps));
    }
    double time(int run) {
        double t = 0;
        check_rc(splash_ensemble_time(e_, run, &t));
        return t;
    }
    bool get_stop_check(int run) {
        int s = 0;
        check_rc(splash_ensemble_stop_check(e_, run, &s));
        return s != 0;
    }
    // resume one run from a state in mesh face order, its time and stop bits
    void set_state(int run, const std::vector<std::vector<double>>& U, double t, int stop_bits = 0) {
        const size_t nf = faces_.at(size_t(run));
        if (U.size() != nf) throw std::invalid_argument("one state row per face");
        std::vector<double> flat(nf * size_t(dim_));
        for (size_t i = 0; i < nf; ++i) {
            if (U[i].size() != size_t(dim_)) throw std::invalid_argument("state rows of dim values");
            for (int k = 0; k < dim_; ++k) flat[i * size_t(dim_) + size_t(k)] = U[i][size_t(k)];
        }
        check_rc(splash_ensemble_set_state(e_, run, flat.data(), t, stop_bits));
    }
    // run's do_step diagnostics sums and light curves (one readout of every run)
    std::vector<double> diagnostics(int run) {
        std::vector<double> out(5);
        check_rc(splash_ensemble_diagnostics(e_, run, out.data()));
        return out;
    }
    std::vector<double> get_light_curves(int run) {
        std::vector<double> out(4);
        check_rc(splash_ensemble_light_curves(e_, run, out.data()));
        return out;
    }
    std::vector<std::vector<double>> get_state(int run) {
        const size_t nf = faces_.at(size_t(run));
        std::vector<double> flat(nf * size_t(dim_));
        check_rc(splash_ensemble_get_state(e_, run, flat.data()));
        std::vector<std::vector<double>> U(nf, std::vector<double>(size_t(dim_)));
        for (size_t i = 0; i < nf; ++i)
            for (int k = 0; k < dim_; ++k) U[i][size_t(k)] = flat[i * size_t(dim_) + size_t(k)];
        return U;
    }

private:
    splash_ensemble* e_ = nullptr;
    int dim_;
    std::vector<size_t> faces_;
};

}  // namespace splash
