// impec_config.hpp -- Impec2D's configuration files (libconfig syntax, read by
// SimParams::SimParams, SimParams.cpp:45-463) without libconfig: a parser for
// the part of the syntax the shipped configurations use -- groups, scalar
// settings (booleans, integers, floats, strings), arrays and lists, and
// "#", "//" and "/* */" comments -- and the reference's mapping of keys onto
// SimParams.  Where the reference exits on a missing or invalid setting,
// ReadConfigFile returns false with the reason.
#pragma once

#include "impec_solvers.hpp"

#include <string>
#include <utility>
#include <vector>

namespace impec_host {

struct CfgValue {
    enum Kind { NONE, BOOL, INT, FLOAT, STRING, ARRAY, LIST, GROUP };
    Kind kind = NONE;
    bool b = false;
    long long i = 0;
    double f = 0.;
    std::string s;
    std::vector<CfgValue> items;                               // ARRAY, LIST
    std::vector<std::pair<std::string, CfgValue>> members;     // GROUP

    const CfgValue *member(const std::string &name) const;     // GROUP lookup (nullptr if absent)
    const CfgValue *path(const std::string &dotted) const;     // "application.general.x_cells"
    // libconfig's lookupValue: false when absent or of another type (a float
    // setting accepts an integer value)
    bool get(const std::string &name, bool &out) const;
    bool get(const std::string &name, int &out) const;
    bool get(const std::string &name, double &out) const;
    bool get(const std::string &name, std::string &out) const;
};

// parses libconfig text into a root group; false with "line N: ..." on error
bool ParseConfig(const std::string &text, CfgValue &root, std::string &err);

// SimParams from a configuration file (SimParams.cpp:45-463); the output
// settings (log_*, results_dir, time lists) are filled as well
bool ReadConfigFile(const std::string &path, SimParams &sp, std::string &err);
bool ReadConfigText(const std::string &text, SimParams &sp, std::string &err);

}  // namespace impec_host
