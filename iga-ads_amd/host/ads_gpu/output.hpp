// Output formats of the reference's post-processing for solutions sampled on the device
// (adsg_sample): gnuplot rows (output/gnuplot.hpp:39-55 with the output manager's
// fixed_format(10, 18), output_manager.hpp:23) and VTK ImageData in ASCII
// (output/vtk.hpp:41-120), plus the uniform axes of output/axis.hpp:15-25.
#pragma once

#include <array>
#include <cstddef>
#include <iomanip>
#include <ostream>
#include <sstream>
#include <vector>

namespace ads_gpu::output {

// util.hpp:28-34: n + 1 points (1 - t) a + t b, t = i / n
inline std::vector<double> linspace(double a, double b, std::size_t n) {
    std::vector<double> xs(n + 1);
    for (std::size_t i = 0; i <= n; ++i) {
        const double t = static_cast<double>(i) / static_cast<double>(n);
        xs[i] = (1 - t) * a + t * b;
    }
    return xs;
}

namespace detail {
inline void num(std::ostream& os, double v) { os << std::setw(18) << v; }
struct fixed10 {
    std::ostream& os;
    std::ios_base::fmtflags flags;
    std::streamsize precision;
    explicit fixed10(std::ostream& o) : os{o}, flags{o.flags()}, precision{o.precision()} {
        os.precision(10);
        os.setf(std::ios_base::fixed, std::ios_base::floatfield);
    }
    ~fixed10() {
        os.flags(flags);
        os.precision(precision);
    }
};
}  // namespace detail

// gnuplot_printer<2>: x outermost; vals[i + nx * j] at (xs[i], ys[j]) (x fastest, as
// adsg_sample writes it)
inline void gnuplot_2d(std::ostream& os, const std::vector<double>& xs, const std::vector<double>& ys,
                       const std::vector<double>& vals) {
    detail::fixed10 guard{os};
    for (std::size_t i = 0; i < xs.size(); ++i)
        for (std::size_t j = 0; j < ys.size(); ++j) {
            detail::num(os, xs[i]);
            detail::num(os, ys[j]);
            detail::num(os, vals[i + xs.size() * j]);
            os << '\n';
        }
}

// output::vtk for one field on an nx x ny x nz grid (x fastest)
inline void vtk_3d(std::ostream& os, const std::array<std::size_t, 3>& n, const std::vector<double>& vals) {
    std::ostringstream ext;
    ext << "0 " << n[0] - 1 << " 0 " << n[1] - 1 << " 0 " << n[2] - 1;
    os << "<?xml version=\"1.0\"?>" << std::endl;
    os << R"(<VTKFile type="ImageData" version="0.1">)" << std::endl;
    os << "  <ImageData WholeExtent=\"" << ext.str() << "\" origin=\"0 0 0\" spacing=\"1 1 1\">" << std::endl;
    os << "    <Piece Extent=\"" << ext.str() << "\">" << std::endl;
    os << "      <PointData Scalars=\"Result\">" << std::endl;
    os << R"(        <DataArray Name="Result"  type="Float32" format="ascii" NumberOfComponents="1">)"
       << std::endl;
    {
        detail::fixed10 guard{os};
        for (double v : vals) {
            detail::num(os, v);
            os << '\n';
        }
    }
    os << "        </DataArray>" << std::endl;
    os << "      </PointData>" << std::endl;
    os << "    </Piece>" << std::endl;
    os << "  </ImageData>" << std::endl;
    os << "</VTKFile>" << std::endl;
}

}  // namespace ads_gpu::output
