// 3D implicit heat with a source written against the GPU host surface, in the shape of the
// reference's implicit example (implicit.hpp:66-111): the directions' M + dt/2 K matrices are
// set up and factored in before(), and every step applies the linearly implicit update
// u += (A_x (x) A_y (x) A_z)^-1 dt (f - K u) (the all-direction form of the reference's
// directional splitting). Initial coefficients and the source come from files.
//
//   implicit_gpu <p> <elements> <steps> <dt> <fp64|fp32> <refine 0|1> <u0.bin> <f.bin> <out.bin>
//
// *.bin: raw little-endian doubles, first index fastest.
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "ads_gpu/simulation.hpp"

class implicit_heat : public ads_gpu::gpu_simulation_3d {
    std::vector<double> u0_, f_;
    bool refine_;

public:
    implicit_heat(const ads_gpu::config_3d& c, std::vector<double> u0, std::vector<double> f,
                  bool refine)
    : gpu_simulation_3d{c}, u0_{std::move(u0)}, f_{std::move(f)}, refine_{refine} {}

    std::vector<double> result() { return state(); }

private:
    void before() override {
        for (int d = 0; d < 3; ++d) set_matrix(d, 1, 1.0, 0.5 * steps.dt);  // implicit.hpp:46-64
        prepare_matrices();
        set_state(u0_);
        set_vector(ADSG_SOURCE, f_);
    }

    void step(int /*iter*/, double /*t*/) override {
        ads_gpu::rhs_form r{steps.dt};
        r.mass = 0.0;
        r.with_source();
        advance_delta(1, r, {1, 1, 1}, refine_);
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
        std::cerr << "usage: implicit_gpu p elements steps dt fp64|fp32 refine u0.bin f.bin out.bin\n";
        return 2;
    }
    const int p = std::atoi(argv[1]), ne = std::atoi(argv[2]), nsteps = std::atoi(argv[3]);
    const double dt = std::atof(argv[4]);
    const int prec = std::strcmp(argv[5], "fp32") == 0 ? ADSG_FP32 : ADSG_FP64;
    const bool refine = std::atoi(argv[6]) != 0;
    try {
        ads_gpu::dim_config dc{p, ne};
        ads_gpu::config_3d c{dc, dc, dc, ads_gpu::timesteps_config{nsteps, dt}, 1, prec};
        implicit_heat sim{c, read_doubles(argv[7]), read_doubles(argv[8]), refine};
        sim.run();
        std::vector<double> out = sim.result();
        std::ofstream o(argv[9], std::ios::binary);
        o.write(reinterpret_cast<const char*>(out.data()), static_cast<std::streamsize>(out.size() * sizeof(double)));
    } catch (const std::exception& e) {
        std::cerr << "implicit_gpu: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
