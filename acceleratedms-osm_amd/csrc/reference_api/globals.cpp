// Shared state and helpers of the reference-compatible API (utils/globals.cpp:3-19).
#include "globals.h"

#include <cstdlib>
#include <stdexcept>

bool verbose = false;

void banner()
{
    cout << "    MS-OSM coherent dedispersion (MI355X)" << endl;
}

void line()
{
    cout << "------------------------------------------------------------" << endl;
}

namespace msosm {

static bool g_throw = false;

void set_throw_on_error(bool on) { g_throw = on; }

void check(int status, const char *what)
{
    if (status == MSOSM_OK) return;
    const std::string msg = std::string(what) + ": " + msosm_last_error();
    if (g_throw) throw std::runtime_error(msg);
    cerr << "Error: " << msg << endl;
    exit(1);
}

} // namespace msosm
