// Probe of the C++ host API's beam filters (dxmclib_amd/include/dxmc/
// beamfilters.hpp), built against the product headers the way a user of the
// reference writes one. Reads commands from stdin, one per line, and prints
// each result as C99 hexadecimal floats (exact):
//
//   xcare <filterAngle> <span> <ramp> <lowWeight> <angle>   (radians)
//       -> the weight of XCareFilter<double> after the four setters
//   bowtie <n> <angle_1 .. angle_n> <weight_1 .. weight_n> <nq> <q_1 .. q_nq>
//       -> line 1: the normalised |angle| weight pairs; line 2: the weights
//          of sampleIntensityWeight at the nq angles
#include "dxmc/beamfilters.hpp"

#include <cstdio>
#include <iostream>
#include <string>
#include <vector>

using namespace dxmc;

int main()
{
    std::string cmd;
    while (std::cin >> cmd) {
        if (cmd == "xcare") {
            double fa, span, ramp, low, angle;
            std::cin >> fa >> span >> ramp >> low >> angle;
            XCareFilter<double> f;
            f.setFilterAngle(fa);
            f.setSpanAngle(span);
            f.setRampAngle(ramp);
            f.setLowWeight(low);
            std::printf("%a\n", f.sampleIntensityWeight(angle));
        } else if (cmd == "bowtie") {
            std::size_t n = 0, nq = 0;
            std::cin >> n;
            std::vector<double> a(n), w(n);
            for (auto& x : a)
                std::cin >> x;
            for (auto& x : w)
                std::cin >> x;
            std::cin >> nq;
            std::vector<double> q(nq);
            for (auto& x : q)
                std::cin >> x;
            BowTieFilter<double> f(a, w);
            for (const auto& p : f.data())
                std::printf("%a %a ", p.first, p.second);
            std::printf("\n");
            for (const double x : q)
                std::printf("%a ", f.sampleIntensityWeight(x));
            std::printf("\n");
        } else {
            std::fprintf(stderr, "unknown command %s\n", cmd.c_str());
            return 2;
        }
    }
    return 0;
}
