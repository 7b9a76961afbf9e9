// sop-gpu <parameter file> [name=value ...]
//
// The reference's command line (src/sop-gpu.cpp:18-88): one argument, the
// parameter file.  Further `name=value` arguments override parameters of
// the file.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sopgpu.h"

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "ERROR: Configuration file should be specified.\n");
        return 1;
    }
    std::vector<std::string> names, values;
    for (int k = 2; k < argc; ++k) {
        const char* eq = std::strchr(argv[k], '=');
        if (!eq) continue;  // the reference ignores further arguments
        names.emplace_back(argv[k], eq - argv[k]);
        values.emplace_back(eq + 1);
    }
    std::vector<const char*> n, v;
    for (size_t k = 0; k < names.size(); ++k) {
        n.push_back(names[k].c_str());
        v.push_back(values[k].c_str());
    }
    const int rc = sop_run_config(argv[1], (int)n.size(), n.data(), v.data());
    if (rc) {
        std::fprintf(stderr, "ERROR: %s\n", sop_last_error());
        return 1;
    }
    return 0;
}
