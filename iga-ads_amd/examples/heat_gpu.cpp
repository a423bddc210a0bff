// Explicit heat problem written against the GPU host surface the way the reference's
// examples are written against simulation_2d / simulation_3d: the problem class overrides
// the simulation_base hooks (heat_3d.hpp:29-47, validation.hpp:65-93) and calls the
// protected helpers. Initial coefficients come from a file instead of an L2 projection.
//
//   heat_gpu <dim 2|3> <p> <elements> <steps> <dt> <dirichlet 0|1> <fp64|fp32> <in.bin> <out.bin>
//
// in.bin / out.bin: raw little-endian doubles, first index fastest.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "ads_gpu/simulation.hpp"

template <typename Base, typename Config>
class heat : public Base {
    std::vector<double> init_;
    bool dirichlet_;

public:
    heat(const Config& c, std::vector<double> init, bool dirichlet)
    : Base{c}, init_{std::move(init)}, dirichlet_{dirichlet} {}

    std::vector<double> result() { return this->state(); }

private:
    void before() override {
        if (dirichlet_)
            for (int d = 0; d < static_cast<int>(this->shape().size()); ++d) {
                this->fix_left(d);
                this->fix_right(d);
            }
        this->prepare_matrices();
        this->set_state(init_);
    }

    void before_step(int /*iter*/, double /*t*/) override { this->swap_state(); }

    void step(int /*iter*/, double /*t*/) override {
        this->compute_rhs(ads_gpu::rhs_form::heat(this->steps.dt));
        if (dirichlet_) this->zero_boundary();
        this->solve();
    }
};

static std::vector<double> read_doubles(const char* path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) throw std::runtime_error(std::string("cannot open ") + path);
    std::streamsize n = f.tellg();
    f.seekg(0);
    std::vector<double> v(static_cast<size_t>(n) / sizeof(double));
    f.read(reinterpret_cast<char*>(v.data()), n);
    return v;
}

int main(int argc, char** argv) {
    if (argc != 10) {
        std::cerr << "usage: heat_gpu dim p elements steps dt dirichlet fp64|fp32 in.bin out.bin\n";
        return 2;
    }
    const int dim = std::atoi(argv[1]), p = std::atoi(argv[2]), ne = std::atoi(argv[3]);
    const int nsteps = std::atoi(argv[4]);
    const double dt = std::atof(argv[5]);
    const bool dirichlet = std::atoi(argv[6]) != 0;
    const int prec = std::strcmp(argv[7], "fp32") == 0 ? ADSG_FP32 : ADSG_FP64;
    try {
        std::vector<double> init = read_doubles(argv[8]);
        ads_gpu::dim_config dc{p, ne};
        ads_gpu::timesteps_config ts{nsteps, dt};
        std::vector<double> out;
        if (dim == 2) {
            ads_gpu::config_2d c{dc, dc, ts, 1, prec};
            heat<ads_gpu::gpu_simulation_2d, ads_gpu::config_2d> sim{c, init, dirichlet};
            sim.run();
            out = sim.result();
        } else {
            ads_gpu::config_3d c{dc, dc, dc, ts, 1, prec};
            heat<ads_gpu::gpu_simulation_3d, ads_gpu::config_3d> sim{c, init, dirichlet};
            sim.run();
            out = sim.result();
        }
        std::ofstream f(argv[9], std::ios::binary);
        f.write(reinterpret_cast<const char*>(out.data()), static_cast<std::streamsize>(out.size() * sizeof(double)));
        std::cout << "heat_gpu: " << out.size() << " coefficients after " << nsteps << " steps\n";
    } catch (const ads_gpu::error& e) {
        std::cerr << "heat_gpu: " << e.what() << " (code " << e.code << ")\n";
        return 1;
    } catch (const std::exception& e) {
        std::cerr << "heat_gpu: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
