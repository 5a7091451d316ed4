// gen_draws.cpp -- prints particle coordinates drawn the way the reference's
// particle generator draws them (tools/UniformGenerator.cpp:16-27,
// ClusteredGaussGenerator.cpp:47-72): libstdc++ engines and distributions,
// three draws per particle passed to std::make_tuple, whose argument
// evaluation order is the compiler's (GCC: right to left).  Output is one
// particle per line, x y z as %a hex floats; used by make_golden.py.
#include <cstdio>
#include <random>
#include <tuple>
#include <vector>

int main()
{
    // uniform: mt19937(926762934), U(0,1) * 10 in a 10^3 box, 8 particles
    {
        std::mt19937 eng(926762934u);
        std::uniform_real_distribution<double> u(0.0, 1.0);
        for (int i = 0; i < 8; ++i) {
            auto t = std::make_tuple(i, 0.0 + u(eng) * 10.0, 0.0 + u(eng) * 10.0, 0.0 + u(eng) * 10.0);
            std::printf("uniform %a %a %a\n", std::get<1>(t), std::get<2>(t), std::get<3>(t));
        }
    }
    // clustered gauss: 3 centres from mt19937(89347587) in a 20^3 box,
    // offsets N(0, 1) from default_random_engine(926762934), 4 per cluster
    {
        std::mt19937 eng(89347587u);
        std::uniform_real_distribution<double> u(0.0, 1.0);
        std::vector<std::tuple<double, double, double>> centres;
        for (int c = 0; c < 3; ++c)
            centres.push_back(std::make_tuple(0.0 + u(eng) * 20.0, 0.0 + u(eng) * 20.0, 0.0 + u(eng) * 20.0));
        std::default_random_engine g(926762934u);
        std::normal_distribution<double> d0{0.0, 1.0}, d1{0.0, 1.0}, d2{0.0, 1.0};
        for (auto &c : centres)
            for (int i = 0; i < 4; ++i) {
                auto t = std::make_tuple(i, std::get<0>(c) + d0(g), std::get<1>(c) + d1(g), std::get<2>(c) + d2(g));
                std::printf("clustered %a %a %a\n", std::get<1>(t), std::get<2>(t), std::get<3>(t));
            }
    }
    // engine known answers fixed by the C++ standard
    {
        std::mt19937 a;
        for (int i = 0; i < 9999; ++i) a();
        std::minstd_rand0 b;
        for (int i = 0; i < 9999; ++i) b();
        std::printf("kat %u %u\n", (unsigned)a(), (unsigned)b());
    }
    return 0;
}
