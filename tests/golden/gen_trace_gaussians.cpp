// Writes the first values that the KPM trace sampling draws: a
// default-constructed std::default_random_engine feeding
// std::normal_distribution<double>(0, 1), one value per line.  Built and run
// by make_trace_gaussians.sh with this image's g++ / libstdc++; the output is
// the fixture trace_gaussians.txt that pins the oracle's restatement of the
// generator (src/locality.cpp:2275-2276, 2301-2303 draw the trace vectors so).
#include <cstdio>
#include <cstdlib>
#include <random>

int main(int argc, char** argv)
{
    const int count = argc > 1 ? std::atoi(argv[1]) : 4000;
    std::default_random_engine generator;
    std::normal_distribution<double> distribution(0.0, 1.0);
    for (int i = 0; i < count; ++i) std::printf("%.17g\n", distribution(generator));
    return 0;
}
