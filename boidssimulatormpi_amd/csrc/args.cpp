// args.cpp — Boost-free restatement of the reference's option handling
// (src/args.cpp:17-153).
//
// Accepted forms: "-N 1024", "-N1024", "--population 1024",
// "--population=1024".  The ini reader follows Boost.PropertyTree's INI
// rules as the reference files use them: [Section] headers, "KEY = value",
// lines starting with ';' or '#' are comments, surrounding blanks trimmed;
// a value that does not convert completely to the requested type is
// treated as absent (get_optional returns none).
//
// Deliberate difference: the reference binds "-i/--initialization" to the
// output file name as well (src/args.cpp:32), so "-i FILE" without "-o"
// overwrites the init file with the output; here "-i" sets only the
// initialization.
#include "args.h"

#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

namespace {

const char* kUsage =
    ":\n"
    "  -h [ --help ]              help\n"
    "  -s [ --setting ] arg       simulation setting file with .ini format.\n"
    "  -o [ --output ] arg        output file.\n"
    "  --velocity-output arg      output file for velocity data.\n"
    "  -p [ --parallel-output ]   output data on several nodes on MPI.\n"
    "  --force-data-output        force vector output on data. (this option is experimental!!!)\n"
    "\n"
    "simulation parameters (have priority over setting file.):\n"
    "  -i [ --initialization ] arg   initialization type or file name. if file name is given, simulation\n"
    "                                will start with last conditon of the file.\n"
    "  -N [ --population ] arg    Number of boids.\n"
    "  -T [ --time-step ] arg     Simulation time step.\n"
    "  -F [ --field-size ] arg    Size of simulation area.\n"
    "  --vmin arg                 minimum velocity\n"
    "  --vmax arg                 maximum velocity\n"
    "  --random-seed arg          random seed\n";

std::string trim(const std::string& s)
{
    size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}

template <typename T>
bool convert(const std::string& text, T* out)
{
    std::istringstream is(text);
    T v;
    is >> v;
    if (is.fail()) return false;
    is >> std::ws;
    if (!is.eof()) return false;
    *out = v;
    return true;
}

template <>
bool convert<std::string>(const std::string& text, std::string* out)
{
    *out = text;
    return true;
}

std::map<std::string, std::string> read_ini(const std::string& path)
{
    std::ifstream f(path);
    if (!f) {
        std::cout << path << ": cannot open file" << std::endl;
        std::exit(-1);
    }
    std::map<std::string, std::string> kv;
    std::string line, section;
    while (std::getline(f, line)) {
        std::string t = trim(line);
        if (t.empty() || t[0] == ';' || t[0] == '#') continue;
        if (t[0] == '[') {
            size_t e = t.find(']');
            section = trim(t.substr(1, e == std::string::npos ? std::string::npos : e - 1));
            continue;
        }
        size_t eq = t.find('=');
        if (eq == std::string::npos) continue;
        kv[section + "." + trim(t.substr(0, eq))] = trim(t.substr(eq + 1));
    }
    return kv;
}

template <typename T>
void from_ini(const std::map<std::string, std::string>& kv, const char* key, T* out)
{
    auto it = kv.find(key);
    if (it != kv.end()) convert(it->second, out);
}

struct Opt {
    const char* lng;
    char shrt;
    bool takes_value;
};

const Opt kOpts[] = {
    {"help", 'h', false},         {"setting", 's', true},     {"output", 'o', true},
    {"velocity-output", 0, true}, {"parallel-output", 'p', false}, {"force-data-output", 0, false},
    {"initialization", 'i', true}, {"population", 'N', true}, {"time-step", 'T', true},
    {"field-size", 'F', true},    {"vmin", 0, true},          {"vmax", 0, true},
    {"random-seed", 0, true},
};

[[noreturn]] void bad(const std::string& msg)
{
    std::cout << msg << std::endl;
    std::exit(-1);
}

}  // namespace

Args::Args(int argc, const char** argv)
{
    std::map<std::string, std::string> given;   // long name -> value ("" for flags)
    for (int k = 1; k < argc; ++k) {
        std::string a = argv[k];
        const Opt* opt = nullptr;
        std::string value;
        bool has_value = false;
        if (a.rfind("--", 0) == 0) {
            std::string name = a.substr(2);
            size_t eq = name.find('=');
            if (eq != std::string::npos) { value = name.substr(eq + 1); name = name.substr(0, eq); has_value = true; }
            for (const Opt& o : kOpts) if (name == o.lng) opt = &o;
            if (!opt) bad("unrecognised option '" + a + "'");
        } else if (a.size() >= 2 && a[0] == '-') {
            for (const Opt& o : kOpts) if (o.shrt && a[1] == o.shrt) opt = &o;
            if (!opt) bad("unrecognised option '" + a + "'");
            if (a.size() > 2) { value = a.substr(2); has_value = true; }
        } else {
            bad("unexpected positional argument '" + a + "'");
        }
        if (opt->takes_value && !has_value) {
            if (k + 1 >= argc) bad(std::string("the required argument for option '--") + opt->lng + "' is missing");
            value = argv[++k];
        }
        if (given.count(opt->lng)) bad(std::string("option '--") + opt->lng + "' cannot be specified more than once");
        given[opt->lng] = value;
    }
    auto has = [&](const char* n) { return given.count(n) != 0; };
    auto conv = [&](const char* n, auto* out) {
        if (!convert(given[n], out)) bad(std::string("the argument ('") + given[n] + "') for option '--" + n + "' is invalid");
    };

    if (has("help")) {
        std::cout << kUsage << std::endl;
        std::exit(0);
    }
    if (has("setting")) setting_filename = given["setting"];
    if (has("output")) output_filename = given["output"];
    if (has("velocity-output")) velocity_output_filename = given["velocity-output"];
    is_parallel_output = false;   // MPI-only in the reference (src/args.cpp:50-55)
    is_force_data_output = has("force-data-output");

    if (has("setting")) {
        auto kv = read_ini(setting_filename);
        from_ini(kv, "Global.TIME_STEP", &time_step);
        from_ini(kv, "Global.POPULATION", &population);
        from_ini(kv, "Global.FIELD_SIZE", &field_size);
        from_ini(kv, "Global.INIT", &initialization);
        from_ini(kv, "Global.RANDOM_SEED", &random_seed);
        from_ini(kv, "Separation.SIGHT_DISTANCE", &separation_sight_distance);
        from_ini(kv, "Separation.SIGHT_ANGLE", &separation_sight_angle);
        from_ini(kv, "Separation.FORCE_COEFFICIENT", &separation_force_coefficient);
        from_ini(kv, "Alignment.SIGHT_DISTANCE", &alignment_sight_distance);
        from_ini(kv, "Alignment.SIGHT_ANGLE", &alignment_sight_angle);
        from_ini(kv, "Alignment.FORCE_COEFFICIENT", &alignment_force_coefficient);
        from_ini(kv, "Cohesion.SIGHT_DISTANCE", &cohesion_sight_distance);
        from_ini(kv, "Cohesion.SIGHT_ANGLE", &cohesion_sight_angle);
        from_ini(kv, "Cohesion.FORCE_COEFFICIENT", &cohesion_force_coefficient);
        from_ini(kv, "Velocity.MAX", &velocity_max);
        from_ini(kv, "Velocity.MIN", &velocity_min);
    }
    if (has("initialization")) initialization = given["initialization"];
    if (has("field-size")) conv("field-size", &field_size);
    if (has("population")) conv("population", &population);
    if (has("time-step")) conv("time-step", &time_step);
    if (has("vmin")) conv("vmin", &velocity_min);
    if (has("vmax")) conv("vmax", &velocity_max);
    if (has("random-seed")) conv("random-seed", &random_seed);
}
