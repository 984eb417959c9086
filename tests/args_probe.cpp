// Test helper: parses argv with the drop-in Args and prints every field.
#include <cstdio>

#include "args.h"

int main(int argc, char** argv)
{
    Args a(argc, (const char**)argv);
    std::printf("setting=%s\noutput=%s\nvelocity_output=%s\nforce=%d\nT=%u\nN=%u\nF=%.17g\n"
                "sep=%.17g,%.17g,%.17g\nali=%.17g,%.17g,%.17g\ncoh=%.17g,%.17g,%.17g\n"
                "vmax=%.17g\nvmin=%.17g\ninit=%s\nseed=%d\n",
                a.setting_filename.c_str(), a.output_filename.c_str(), a.velocity_output_filename.c_str(),
                (int)a.is_force_data_output, a.time_step, a.population, a.field_size,
                a.separation_sight_distance, a.separation_sight_angle, a.separation_force_coefficient,
                a.alignment_sight_distance, a.alignment_sight_angle, a.alignment_force_coefficient,
                a.cohesion_sight_distance, a.cohesion_sight_angle, a.cohesion_force_coefficient,
                a.velocity_max, a.velocity_min, a.initialization.c_str(), a.random_seed);
    return 0;
}
